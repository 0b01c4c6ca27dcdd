"""Host side of the product on the CPU: the C ABI library loads and exports
every entry point include/dfsph.h declares, its host utilities agree with the
oracle bit for bit, and argument errors are reported as status codes."""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "dfsph.h")


def _declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(dfsph_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol(dfsph_mod):
    names = _declared_functions()
    assert len(names) >= 14
    lib = C.CDLL(dfsph_mod.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(dfsph_mod.EXPORTED)


def _gpu_present():
    return os.path.exists("/dev/kfd") and any(os.path.exists(f"/dev/dri/renderD{128 + i}") for i in range(8))


@pytest.mark.skipif(_gpu_present(), reason="a GPU is present; covered by the gpu tests")
def test_create_without_gpu_fails_with_device_error(dfsph_mod):
    D = dfsph_mod
    with pytest.raises(D.DfsphError) as e:
        D.Sim(D.make_params(0.01, (0, 0, 0), (1, 1, 1), 100))
    assert e.value.code == D.E_DEVICE


@pytest.mark.parametrize("box,r", [((2.0, 1.0, 1.0), 0.01), ((2.6, 1.3, 1.3), 0.005),
                                   ((3.2, 0.8, 1.6), 0.0025), ((1.7, 0.9, 0.35), 0.013)])
def test_derived_parameters_match_oracle(dfsph_mod, oracle_mod, box, r):
    D, O = dfsph_mod, oracle_mod
    d = D.derive(D.make_params(r, (0, 0, 0), box, 10))
    od = O.derive(O.make_config(r, (0, 0, 0), box))
    for f in ("h", "cell", "mass", "knorm", "gnorm", "relax", "max_axis", "total_cells"):
        assert getattr(d, f) == getattr(od, f), f
    assert list(d.axes) == list(od.axes)
    # FINE grid: cell edge >= h, enough cells to cover the box
    assert d.fine_cell >= d.h
    assert all(a * d.fine_cell >= e for a, e in zip(d.fine_axes, box))


def test_pair_thresholds_are_exact(dfsph_mod):
    """r^2 < r2_q1 <=> fl(fl(sqrt(r^2))/h) < 1, checked at and around the threshold."""
    D = dfsph_mod
    for r in (0.01, 0.005, 0.0025, 0.013):
        d = D.derive(D.make_params(r, (0, 0, 0), (1, 1, 1), 10))
        h = np.float32(d.h)
        t = np.float32(d.r2_q1)
        for v in (np.nextafter(t, np.float32(0)), t, np.nextafter(t, np.float32(1))):
            assert (v < t) == bool(np.sqrt(v) / h < np.float32(1.0))
        e = np.float32(1e-7)
        for thr, pred in ((d.r2_eps_gt, lambda v: np.sqrt(v) > e), (d.r2_eps_lt, lambda v: not (np.sqrt(v) < e))):
            thr = np.float32(thr)
            assert pred(thr) and not pred(np.nextafter(thr, np.float32(0)))


@pytest.mark.parametrize("box,r,cap", [((2.0, 1.0, 1.0), 0.01, 0), ((2.6, 1.3, 1.3), 0.005, 0),
                                       ((0.9, 0.7, 0.3), 0.013, 0), ((3.2, 0.8, 1.6), 0.0025, 512000)])
def test_boundary_samples_match_oracle(dfsph_mod, oracle_mod, box, r, cap):
    a = dfsph_mod.boundary_samples((0, 0, 0), box, r, cap)
    b = oracle_mod.boundary_samples((0, 0, 0), box, r, cap)
    assert a.shape == b.shape and np.array_equal(a, b)
    if cap:
        assert a.shape[0] == cap


def _splitmix64(z):
    z = z + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def test_scene_block_is_the_documented_lattice(dfsph_mod):
    """origin + (i + 0.5) 2r (shaders/initial_fill.comp.glsl:37-41) plus
    (2u - 1) * jitter * r, u = (splitmix64((seed << 40) ^ (3 id + axis)) >> 40) / 2^24."""
    nx, ny, nz, r, seed, jit, id0 = 7, 5, 3, 0.01, 11, 0.1, 40
    org = (0.02, 0.03, 0.04)
    x = dfsph_mod.scene_block(nx, ny, nz, r, org, seed, jit, id0)
    gid = np.arange(nx * ny * nz, dtype=np.uint64)
    idx = [gid % np.uint64(nx), (gid // np.uint64(nx)) % np.uint64(ny), gid // np.uint64(nx * ny)]
    with np.errstate(over="ignore"):
        for a in range(3):
            key = (np.uint64(seed) << np.uint64(40)) ^ ((gid + np.uint64(id0)) * np.uint64(3) + np.uint64(a))
            u = (_splitmix64(key) >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -24)
            j = (np.float32(2) * u - np.float32(1)) * (np.float32(jit) * np.float32(r))
            ref = (np.float32(org[a]) + (idx[a].astype(np.float32) + np.float32(0.5)) * (np.float32(2) * np.float32(r))) + j
            assert np.array_equal(x[:, a], ref)


def test_argument_errors_are_status_codes(dfsph_mod):
    D = dfsph_mod
    bad = D.make_params(0.01, (0, 0, 0), (1, 1, 1), 10)
    bad.abi_version = 99
    with pytest.raises(D.DfsphError) as e:
        D.derive(bad)
    assert e.value.code == D.E_INVALID and "abi_version" in str(e.value)
    for mutate in (lambda p: setattr(p, "capacity", 0), lambda p: setattr(p, "particle_radius", -1.0),
                   lambda p: p.domain_max.__setitem__(1, -1.0), lambda p: setattr(p, "grid", 7),
                   lambda p: setattr(p, "block_size", 100)):
        p = D.make_params(0.01, (0, 0, 0), (1, 1, 1), 10)
        mutate(p)
        with pytest.raises(D.DfsphError) as e:
            D.derive(p)
        assert e.value.code == D.E_INVALID


def test_default_policy_mirrors_reference_constants(dfsph_mod):
    """main.cpp:115-122, :259-261."""
    D = dfsph_mod
    p = D.default_policy(D.POLICY_REFERENCE)
    assert (p.min_iter_div, p.max_iter_div, p.min_iter_density, p.max_iter_density) == (1, 100, 2, 100)
    assert p.div_enabled == 1 and p.warm_start == D.WARM_ZERO and p.boundary_psi == -1
