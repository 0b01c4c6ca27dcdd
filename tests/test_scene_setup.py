"""Scene presets, initial fill and emitter placement on the host (no GPU).

The library's preset fill (dfsph_preset_fill) is checked bit for bit against
the oracle's restatement of seedInitialFill + initial_fill
(main.cpp:2266-2294, shaders/initial_fill.comp.glsl:22-65) for every preset
and capacity cut, and both against properties the formulas fix: the brick is
the lattice bmin + (ijk + 1/2) 2r inside the brick, the droplet lies within
its radius, flags mark droplet particles.  The emitter oracle
(shaders/apply_emitter.comp.glsl:29-65) is pinned by what the sampler must
satisfy: every point within the emitter radius, about pi/6 of the raw cube
samples already inside the ball, the PCG hash injective over 2^20 inputs.
The hash constants themselves stay unpinned (no reference output exists).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle as O


def test_presets(dfsph_mod):
    D = dfsph_mod
    names = [D.preset(k).name.decode() for k in range(D.PRESET_COUNT)]
    assert names == ["Dam-Break", "Central-Fountain", "Droplet-Impact", "Pour-from-Source"]
    p = D.preset(2)
    assert p.add_droplet == 1 and abs(p.droplet_radius - 0.15) < 1e-7
    assert D.preset(3).source_shape == 2 and D.preset(1).source_shape == 1
    for bad in (-1, D.PRESET_COUNT):
        with pytest.raises(D.DfsphError):
            D.preset(bad)


@pytest.mark.parametrize("k", range(4))
@pytest.mark.parametrize("r", [0.01, 0.02])
def test_fill_matches_oracle(dfsph_mod, k, r):
    D = dfsph_mod
    p = D.preset(k)
    drop = (list(p.droplet_center) + [p.droplet_radius] + list(p.droplet_velocity)) if p.add_droplet else None
    full = D.preset_fill(p, r, 1 << 26)
    xo, vo, fo, nbo = O.initial_fill(list(p.brick_min), list(p.brick_max), r, 1 << 26, droplet=drop)
    x, v, f, nb = full
    assert nb == nbo and np.array_equal(x, xo) and np.array_equal(v, vo) and np.array_equal(f, fo)
    # capacity cuts: inside the brick, and inside the droplet
    for cap in (max(nb - 5, 1), nb + 10):
        xc, vc, fc, nbc = D.preset_fill(p, r, cap)
        assert xc.shape[0] <= cap and np.array_equal(xc, x[:xc.shape[0]])
        assert nbc == min(nb, cap)


@pytest.mark.parametrize("k", range(4))
def test_fill_properties(dfsph_mod, k):
    D = dfsph_mod
    r, sp = 0.01, np.float32(0.02)
    p = D.preset(k)
    x, v, f, nb = D.preset_fill(p, r, 1 << 26)
    bmin, bmax = np.array(p.brick_min, np.float32), np.array(p.brick_max, np.float32)
    b = x[:nb]
    assert np.all(f[:nb] == 0) and np.all(v[:nb] == 0)
    ext = bmax - bmin
    if np.all(ext > 2 * sp):
        # lattice: first point at bmin + sp/2, unit steps of sp, inside the brick
        assert np.allclose(b.min(0), bmin + sp / 2, atol=1e-6)
        assert np.all(b <= bmax + 1e-6)
        for a in range(3):
            u = np.unique(b[:, a])
            if u.size > 1:
                assert np.allclose(np.diff(u), sp, atol=2e-6)
        # one more lattice point along any axis would leave the brick
        n_axis = [np.unique(b[:, a]).size for a in range(3)]
        assert math.prod(n_axis) == nb
        for a in range(3):
            assert bmin[a] + (n_axis[a] + 0.5) * sp > bmax[a] - 1e-6
    d = x[nb:]
    if p.add_droplet:
        c = np.array(p.droplet_center, np.float32)
        dr = int(p.droplet_radius / 0.02)
        assert d.shape[0] == (4 * dr ** 3) // 3
        assert np.all(np.linalg.norm(d - c, axis=1) <= p.droplet_radius * (1 + 1e-6))
        assert np.all(f[nb:] == 2) and np.allclose(v[nb:], np.array(p.droplet_velocity, np.float32))
    else:
        assert d.shape[0] == 0


def test_fill_rejects_bad_input(dfsph_mod):
    D = dfsph_mod
    p = D.preset(0)
    with pytest.raises(D.DfsphError):
        D.preset_fill(p, 0.0, 100)
    p.brick_max[0] = p.brick_min[0] - 1.0
    with pytest.raises(D.DfsphError):
        D.preset_fill(p, 0.01, 100)


def test_emitter_oracle_properties():
    em = np.array([[0.0, 1.0, 0.0, 0.1, 0.0, 5.0, 0.0, 5000.0],
                   [0.5, -0.2, 0.3, 0.05, 1.0, 0.0, 0.0, 100.0]], np.float32)
    k = 20000
    x, v = O.emit(em, 1234, k)
    e = np.arange(k) % 2
    c, rad = em[e, 0:3], em[e, 3]
    dist = np.linalg.norm(x.astype(np.float64) - c, axis=1)
    assert np.all(dist <= rad * (1 + 1e-5))
    assert np.array_equal(v, em[e, 4:7])
    # about pi/6 of the cube samples fall inside the ball (not pulled to its surface)
    inside = np.mean(dist < rad * (1 - 1e-4))
    assert abs(inside - math.pi / 6) < 0.02, inside
    assert np.all(np.abs((x - c).mean(0) / rad.mean()) < 0.02)
    # a point depends on its slot and emitter index (g % count restarts with each batch)
    x2, _ = O.emit(em, 1234 + 10, 6)
    assert np.array_equal(x2, x[10:16])           # same slots, same emitter parity
    x2, _ = O.emit(em, 1234 + 11, 6)
    assert not np.any(np.all(x2 == x[11:17], axis=1))   # same slots, other emitters
    x3, _ = O.emit(em, 1234, 16)
    assert np.array_equal(x3, x[:16])


def test_pcg_hash_injective():
    # unit_ball_sample(seed) takes three consecutive hashes; samples of 2^18
    # distinct seeds must all differ (the hash is a bijection of u32).
    em = np.array([[0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0]], np.float32)
    x, _ = O.emit(em, 0, 1 << 18)
    assert np.unique(x, axis=0).shape[0] == x.shape[0]
