"""Pins of the render oracle (oracle/render_oracle.cpp), each against an
independent float64 evaluation of the reference shaders' definitions:

* a single sphere sprite straight ahead: the covered pixels (centre inside the
  gl_PointSize square and its inscribed disc) and each one's impostor NDC z
  (particle_sprite.frag.glsl:29-40) from the closed form of the perspective
  projection;
* the depth pass keeps the last draw (no depth test in the reference): with two
  sprites over one pixel the higher index wins even when it is farther;
* the thickness profile tpp exp(-2 r^2) / z per pixel (thickness.frag.glsl),
  additive in draw order (two coincident particles give exactly twice one);
* point clipping: a point behind the camera or outside the frustum draws nothing;
* the bilateral filter (bilateral_smooth.comp.glsl): a Gaussian-weighted mean
  along x when the depth weight is flat, edges kept when it is sharp, sky
  pixels untouched, and rows left alone unless the alternate mode filters y;
* the composite (composite.frag.glsl): sky colour per view ray, and on a flat
  sheet facing the camera the whole Fresnel / refraction / reflection /
  Beer-Lambert / roughness chain per pixel.
CPU only."""
from __future__ import annotations

import math

import numpy as np
import pytest

NEAR, FAR = 0.05, 100.0
HOR, ZEN, TINT = np.array([0.85, 0.90, 0.95]), np.array([0.40, 0.55, 0.85]), np.array([0.10, 0.30, 0.50])


def camera(eye, target, fov_deg, w, h):
    """glm::lookAt and glm::perspective (depth -1..1), y flipped, float64 -> float32."""
    eye, target = np.asarray(eye, float), np.asarray(target, float)
    f = target - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, [0.0, 1.0, 0.0])
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    view = np.eye(4)
    view[0, :3], view[1, :3], view[2, :3] = s, u, -f
    view[:3, 3] = [-s @ eye, -u @ eye, f @ eye]
    th = math.tan(math.radians(fov_deg) / 2)
    proj = np.zeros((4, 4))
    proj[0, 0] = 1 / (w / h * th)
    proj[1, 1] = -1 / th
    proj[2, 2] = -(FAR + NEAR) / (FAR - NEAR)
    proj[3, 2] = -1.0
    proj[2, 3] = -2 * FAR * NEAR / (FAR - NEAR)
    # column-major flattening = transpose then row-major ravel
    return view.T.astype(np.float32).ravel(), proj.T.astype(np.float32).ravel(), h / (2 * th)


def ndc_z(view_z):
    return ((-(FAR + NEAR) / (FAR - NEAR)) * view_z - 2 * FAR * NEAR / (FAR - NEAR)) / (-view_z)


def sky(d):
    t = np.clip(d[..., 1] * 0.5 + 0.5, 0, 1)
    e = np.clip(t / 0.7, 0, 1)
    s = e * e * (3 - 2 * e)
    return HOR * (1 - s)[..., None] + ZEN * s[..., None]


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


W = H = 64
EYE, TGT = (0.0, 0.0, 3.0), (0.0, 0.0, 0.0)


def _render(O, x, radius=0.2, **kw):
    view, proj, ps = camera(EYE, TGT, 60.0, W, H)
    return O.render(np.asarray(x, np.float32), view, proj, EYE, W, H, radius, ps, **kw)


def _sprite_expect(center_z, radius, ps):
    """Covered mask, uv and depth of a sphere on the view axis at view z."""
    size = min(max(2 * radius * ps / -center_z, 2.0), 1023.0)
    px = np.arange(W) + 0.5
    sc = 0.5 + (px - W / 2) / size
    u = sc * 2 - 1
    uu, vv = np.meshgrid(u, u)          # rows: y (t), columns: x (s)
    inside = (sc >= 0) & (sc < 1)
    r2 = uu ** 2 + vv ** 2
    mask = inside[None, :] & inside[:, None] & (r2 <= 1)
    vz = center_z + np.sqrt(np.clip(1 - r2, 0, None)) * radius
    return mask, ndc_z(vz), r2, size


def test_single_sphere_coverage_and_depth(O):
    _, _, ps = camera(EYE, TGT, 60.0, W, H)
    out = _render(O, [[0.0, 0.0, 0.0]])
    mask, dz, _, size = _sprite_expect(-3.0, 0.2, ps)
    got = out["depth"] < 1.0
    assert got.sum() > 30 and np.array_equal(got, mask)
    assert np.allclose(out["depth"][mask], dz[mask], rtol=2e-6, atol=0)
    # the impostor bulges toward the camera: the smallest NDC z at the centre
    c = np.unravel_index(np.argmin(np.where(mask, out["depth"], 2.0)), mask.shape)
    assert abs(c[0] - (H / 2 - 0.5)) <= 0.5 and abs(c[1] - (W / 2 - 0.5)) <= 0.5


def test_last_draw_wins_without_depth_test(O):
    near, far = [0.0, 0.0, 0.5], [0.0, 0.0, 0.0]
    a = _render(O, [near, far])["depth"]
    b = _render(O, [far, near])["depth"]
    only_far = _render(O, [far])["depth"]
    only_near = _render(O, [near])["depth"]
    both = (only_far < 1) & (only_near < 1)
    assert both.sum() > 20
    assert np.array_equal(a[both], only_far[both])     # index 1 drawn last, though farther
    assert np.array_equal(b[both], only_near[both])
    assert np.all(a[both] > only_near[both])           # the farther sphere's larger NDC z


def test_thickness_profile_and_additivity(O):
    _, _, ps = camera(EYE, TGT, 60.0, W, H)
    one = _render(O, [[0.0, 0.0, 0.0]], tpp=0.005)["thickness"]
    mask, _, r2, _ = _sprite_expect(-3.0, 0.2, ps)
    want = 0.005 * np.exp(-2 * r2) / 3.0
    assert np.array_equal(one > 0, mask)
    assert np.allclose(one[mask], want[mask], rtol=2e-6, atol=0)
    two = _render(O, [[0.0, 0.0, 0.0]] * 2, tpp=0.005)["thickness"]
    assert np.array_equal(two, one + one)


def test_points_outside_the_clip_volume_draw_nothing(O):
    # behind the camera, and just past the right frustum edge (its sprite would reach in)
    edge = 3.0 * math.tan(math.radians(30.0)) * 1.001
    for x in ([0.0, 0.0, 5.0], [edge, 0.0, 0.0]):
        out = _render(O, [x], radius=0.3)
        assert np.all(out["depth"] == 1.0) and np.all(out["thickness"] == 0.0), x
    inside = _render(O, [[edge * 0.99, 0.0, 0.0]], radius=0.3)
    assert (inside["depth"] < 1).sum() > 0


def test_bilateral_along_x(O):
    rng = np.random.default_rng(1)
    # flat depth weight (sigma_d huge): a Gaussian-weighted mean over the row window
    d = rng.uniform(0.2, 0.8, (8, 40)).astype(np.float32)
    got = O.render_bilateral(d, iterations=1, sigma_s=2.5, sigma_d=1e3)
    rad = math.ceil(2 * 2.5)
    want = np.zeros_like(d, dtype=float)
    for y in range(8):
        for x in range(40):
            idx = np.arange(max(0, x - rad), min(40, x + rad + 1))
            w = np.exp(-((idx - x) ** 2) / (2 * 2.5 ** 2))
            want[y, x] = (w * d[y, idx]).sum() / w.sum()
    assert np.allclose(got, want, rtol=0, atol=2e-7)
    # rows that are constant stay put along x; the alternate mode filters y too
    rows = np.repeat(np.linspace(0.3, 0.6, 12, dtype=np.float32)[:, None], 20, axis=1)
    assert np.abs(O.render_bilateral(rows, iterations=4, sigma_d=1.0) - rows).max() <= 4e-7   # rounding only
    assert np.abs(O.render_bilateral(rows, iterations=4, sigma_d=1.0, alternate=True) - rows).max() > 1e-3
    # a sharp depth weight keeps a step; sky pixels neither change nor contribute
    step = np.where(np.arange(30) < 15, 0.3, 0.6).astype(np.float32)[None, :].repeat(4, 0)
    step[:, 20] = 1.0
    out = O.render_bilateral(step, iterations=3, sigma_d=0.01)
    assert np.all(out[:, 20] == 1.0)
    keep = np.arange(30) != 20
    assert np.abs(out[:, keep] - step[:, keep]).max() <= 4e-7


def test_composite_sky_and_flat_sheet(O):
    view, proj, _ = camera(EYE, TGT, 60.0, W, H)
    th = math.tan(math.radians(30.0))
    px = (np.arange(W) + 0.5) / W * 2 - 1
    xx, yy = np.meshgrid(px, px)
    # sky: view ray of each pixel (far plane), y flipped
    rays = np.stack([xx * th, -yy * th, -np.ones_like(xx)], -1)
    rays /= np.linalg.norm(rays, axis=-1, keepdims=True)
    rgba = O.render_composite(np.ones((H, W), np.float32), np.zeros((H, W), np.float32), view, proj, EYE)
    assert np.allclose(rgba[..., :3], sky(rays), atol=2e-6)
    assert np.all(rgba[..., 3] == 1.0)
    # a sheet at world z = 1 facing the camera, constant thickness, roughness
    T, scale, rough, expo = 0.7, 3.0, 0.2, 1.5
    depth = np.full((H, W), ndc_z(-2.0), np.float32)
    rgba = O.render_composite(depth, np.full((H, W), T, np.float32), view, proj, EYE, rough=rough,
                              exposure=expo, absorption=0.005, thick_scale=scale)
    P = np.stack([xx * th * 2.0, -yy * th * 2.0, np.ones_like(xx)], -1)
    V = np.asarray(EYE) - P
    V /= np.linalg.norm(V, axis=-1, keepdims=True)
    N = np.array([0.0, 0.0, 1.0])
    cos = np.clip(V @ N, 0, None)
    F = 0.02 + 0.98 * np.clip(1 - cos, 0, 1) ** 5
    I = -V
    ni = I @ N
    eta = 1 / 1.33
    k = 1 - eta * eta * (1 - ni * ni)
    refr = eta * I - (eta * ni + np.sqrt(np.clip(k, 0, None)))[..., None] * N
    refl = I - 2 * ni[..., None] * N
    trans = np.exp(-0.005 * T * scale * (1 - TINT))
    body = sky(refr) * trans + TINT * (1 - trans)
    col = body * (1 - F[..., None]) + sky(refl) * F[..., None]
    col = col * (1 - rough * 0.3) + (sky(refl) * (1 - rough) + ZEN * rough) * rough * 0.3
    assert np.allclose(rgba[..., :3], col * expo, atol=2e-4)
