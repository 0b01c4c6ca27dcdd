"""ctypes binding of the CPU oracle (oracle/dfsph_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker -- never as the thing
measured or shipped.  The functions restated by the C++ file cite the
reference's files line by line; this module only marshals arrays.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

F_X, F_V, F_RHO, F_AOR2, F_RHO_ADV, F_RHO_CHANGE = 0, 1, 2, 3, 4, 5
F_P_DIV, F_P_DENSITY, F_ACCEL, F_ALPHA, F_NBR, F_NBR_B = 6, 7, 8, 9, 10, 11
F_KEYS, F_SORTED, F_CELL_STARTS, F_BX, F_BV = 12, 13, 14, 15, 16
F_WARM_DIV, F_WARM_DENSITY, F_GF, F_GB = 17, 18, 19, 20
F_RES_DIV, F_RES_DENSITY = 22, 23   # per-particle residuals of each solve's last checked iterate

_VEC_FIELDS = {F_X, F_V, F_ACCEL, F_BX, F_GF, F_GB}
_BOUNDARY_FIELDS = {F_BX, F_BV}


class OracleConfig(C.Structure):
    _fields_ = [
        ("r", C.c_float), ("rho0", C.c_float), ("gravity", C.c_float * 3),
        ("viscosity", C.c_float), ("cohesion", C.c_float), ("vorticity", C.c_float),
        ("dmin", C.c_float * 3), ("dmax", C.c_float * 3),
        ("boundary", C.c_int32), ("boundary_cap", C.c_uint32),
    ]


class OracleDerived(C.Structure):
    _fields_ = [
        ("h", C.c_float), ("cell", C.c_float), ("mass", C.c_float), ("knorm", C.c_float),
        ("gnorm", C.c_float), ("relax", C.c_float), ("axes", C.c_uint32 * 3),
        ("max_axis", C.c_uint32), ("total_cells", C.c_uint32),
    ]


class OraclePolicy(C.Structure):
    _fields_ = [
        ("mode", C.c_int32), ("warm", C.c_int32), ("div_enabled", C.c_int32),
        ("min_iter_div", C.c_int32), ("max_iter_div", C.c_int32),
        ("min_iter_density", C.c_int32), ("max_iter_density", C.c_int32),
        ("tol_div", C.c_float), ("tol_density", C.c_float),
        ("cfl", C.c_float), ("dt_max", C.c_float), ("boundary_psi", C.c_int32),
        ("residual_bound", C.c_float),
    ]


class OracleStats(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("iters_div", C.c_int32), ("iters_density", C.c_int32),
        ("converged_div", C.c_int32), ("converged_density", C.c_int32),
        ("err_div", C.c_double), ("err_density", C.c_double),
        ("max_speed", C.c_double), ("nonfinite", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    srcs = [os.path.join(HERE, f) for f in ("dfsph_oracle.cpp", "render_oracle.cpp")]
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < max(map(os.path.getmtime, srcs)):
        subprocess.run(["make", "-s", "-C", HERE, os.path.join(HERE, "liboracle.so")], check=True)
    return LIB_PATH


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.oracle_create.restype = P
        L.oracle_create.argtypes = [C.POINTER(OracleConfig), C.c_int, C.c_uint32, P, P]
        L.oracle_destroy.argtypes = [P]
        L.oracle_step.restype = C.c_int
        L.oracle_step.argtypes = [P, C.c_float, C.POINTER(OraclePolicy), C.POINTER(OracleStats)]
        L.oracle_get.argtypes = [P, C.c_int, P]
        L.oracle_get_u32.argtypes = [P, C.c_int, P]
        L.oracle_set.argtypes = [P, C.c_int, P]
        L.oracle_count.restype = C.c_uint32
        L.oracle_count.argtypes = [P, C.c_int]
        L.oracle_sim_derived.argtypes = [P, C.POINTER(OracleDerived)]
        L.oracle_derive.argtypes = [C.POINTER(OracleConfig), C.POINTER(OracleDerived)]
        L.oracle_morton3.restype = C.c_uint32
        L.oracle_morton3.argtypes = [C.c_uint32] * 3
        L.oracle_kernel_f32.argtypes = [C.c_float, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.oracle_kernel_f64.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.oracle_boundary_samples.restype = C.c_uint32
        L.oracle_boundary_samples.argtypes = [P, P, C.c_float, C.c_uint32, P, C.c_uint32]
        L.oracle_hash.argtypes = [C.POINTER(OracleConfig), C.c_uint32, P, P, P, P]
        L.oracle_initial_fill.restype = C.c_uint32
        L.oracle_initial_fill.argtypes = [P, C.c_int, P, C.c_float, C.c_uint32, P, P, P, P]
        L.oracle_emit.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32, P, P]
        L.oracle_prepare.argtypes = [P, C.c_int]
        L.oracle_apply_A.argtypes = [P, C.c_int, C.c_double, C.c_int, P, P, P]
        L.oracle_source.argtypes = [P, C.c_int, C.c_double]
        L.oracle_forces.argtypes = [P, C.c_double]
        L.oracle_render.argtypes = [P, C.c_uint32, P, P, P, C.c_uint32, C.c_uint32, C.c_float, C.c_float,
                                    C.c_float, C.c_int, C.c_float, C.c_float, C.c_int, P, C.c_float, P, P,
                                    C.c_float, C.c_float, C.c_float, C.c_float, P, P, P, P]
        L.oracle_render_bilateral.argtypes = [P, C.c_uint32, C.c_uint32, C.c_int, C.c_float, C.c_float, C.c_int, P]
        L.oracle_render_composite.argtypes = [P, P, C.c_uint32, C.c_uint32, P, P, P, P, C.c_float, P, P,
                                              C.c_float, C.c_float, C.c_float, C.c_float, P]
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_max_threads.restype = C.c_int
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def make_config(r, dmin, dmax, rho0=1000.0, gravity=(0.0, -9.81, 0.0), viscosity=0.01,
                cohesion=0.0, vorticity=0.0, boundary=True, boundary_cap=0) -> OracleConfig:
    c = OracleConfig()
    c.r, c.rho0 = r, rho0
    c.gravity[:] = list(gravity)
    c.viscosity, c.cohesion, c.vorticity = viscosity, cohesion, vorticity
    c.dmin[:] = list(dmin)
    c.dmax[:] = list(dmax)
    c.boundary = 1 if boundary else 0
    c.boundary_cap = boundary_cap
    return c


def derive(cfg: OracleConfig) -> OracleDerived:
    d = OracleDerived()
    lib().oracle_derive(C.byref(cfg), C.byref(d))
    return d


def policy(mode="reference", warm="zero", div_enabled=True, min_iter_div=1, max_iter_div=100,
           min_iter_density=2, max_iter_density=100, tol_div=1e-3, tol_density=1e-3,
           cfl=0.4, dt_max=1e-3, boundary_psi=None, residual_bound=0.0) -> OraclePolicy:
    p = OraclePolicy()
    p.mode = {"reference": 0, "converged": 1}[mode]
    p.warm = {"zero": 0, "previous": 1, "external": 2}[warm]
    p.div_enabled = 1 if div_enabled else 0
    p.min_iter_div, p.max_iter_div = min_iter_div, max_iter_div
    p.min_iter_density, p.max_iter_density = min_iter_density, max_iter_density
    p.tol_div, p.tol_density, p.cfl, p.dt_max = tol_div, tol_density, cfl, dt_max
    if boundary_psi is None:
        boundary_psi = 1 if mode == "converged" else 0
    p.boundary_psi = boundary_psi
    p.residual_bound = residual_bound
    return p


def set_threads(n: int) -> None:
    """OpenMP threads of the oracle's loops."""
    lib().oracle_set_threads(n)


def max_threads() -> int:
    return int(lib().oracle_max_threads())


def kernel_f32(q: float, h: float):
    W, g = C.c_float(), C.c_float()
    lib().oracle_kernel_f32(q, h, C.byref(W), C.byref(g))
    return W.value, g.value


def kernel_f64(q: float, h: float):
    W, g = C.c_double(), C.c_double()
    lib().oracle_kernel_f64(q, h, C.byref(W), C.byref(g))
    return W.value, g.value


def morton3(x: int, y: int, z: int) -> int:
    return int(lib().oracle_morton3(x, y, z))


def boundary_samples(dmin, dmax, r, cap=0) -> np.ndarray:
    lo = np.asarray(dmin, np.float32)
    hi = np.asarray(dmax, np.float32)
    n = lib().oracle_boundary_samples(_ptr(lo), _ptr(hi), r, cap, None, 0)
    out = np.zeros((n, 3), np.float32)
    lib().oracle_boundary_samples(_ptr(lo), _ptr(hi), r, cap, _ptr(out), n)
    return out


def initial_fill(brick_min, brick_max, r, capacity, droplet=None):
    """seedInitialFill + initial_fill: droplet = (centre xyz, radius, velocity xyz)
    or None.  Returns (x, v, flags, brick_count)."""
    brick = np.asarray(list(brick_min) + list(brick_max), np.float32)
    dr = np.asarray(droplet if droplet is not None else [0.0] * 7, np.float32)
    nb = C.c_uint32()
    n = lib().oracle_initial_fill(_ptr(brick), 1 if droplet is not None else 0, _ptr(dr), r, capacity,
                                  None, None, None, C.byref(nb))
    x = np.zeros((n, 3), np.float32)
    v = np.zeros((n, 3), np.float32)
    f = np.zeros(n, np.uint32)
    lib().oracle_initial_fill(_ptr(brick), 1 if droplet is not None else 0, _ptr(dr), r, capacity,
                              _ptr(x), _ptr(v), _ptr(f), C.byref(nb))
    return x, v, f, nb.value


def emit(emitters, n0: int, k: int):
    """Emitter placement for slots n0 .. n0+k-1; emitters = rows (pos xyz, radius,
    velocity xyz, rate)."""
    em = np.ascontiguousarray(emitters, np.float32).reshape(-1, 8)
    x = np.zeros((k, 3), np.float32)
    v = np.zeros((k, 3), np.float32)
    lib().oracle_emit(_ptr(em), em.shape[0], n0, k, _ptr(x), _ptr(v))
    return x, v


def hash_reference(cfg: OracleConfig, x: np.ndarray):
    x = np.ascontiguousarray(x, np.float32)
    n = x.shape[0]
    d = derive(cfg)
    keys = np.zeros(n, np.uint32)
    sorted_ = np.zeros(n, np.uint32)
    starts = np.zeros(d.total_cells + 1, np.uint32)
    lib().oracle_hash(C.byref(cfg), n, _ptr(x), _ptr(keys), _ptr(sorted_), _ptr(starts))
    return keys, sorted_, starts


@dataclass
class OracleSim:
    """Stateful oracle: state in original-index order, as the reference keeps it."""
    cfg: OracleConfig
    fp64: bool
    handle: int

    @classmethod
    def create(cls, cfg: OracleConfig, x: np.ndarray, v: np.ndarray | None = None, fp64=False):
        x = np.ascontiguousarray(x, np.float32)
        vv = None if v is None else np.ascontiguousarray(v, np.float32)
        h = lib().oracle_create(C.byref(cfg), 1 if fp64 else 0, x.shape[0], _ptr(x),
                                None if vv is None else _ptr(vv))
        if not h:
            raise RuntimeError("oracle_create failed")
        sim = cls(cfg, fp64, h)
        sim._keep = (x, vv)
        return sim

    def __del__(self):
        if getattr(self, "handle", None):
            lib().oracle_destroy(self.handle)
            self.handle = None

    @property
    def n(self) -> int:
        return int(lib().oracle_count(self.handle, 0))

    @property
    def nb(self) -> int:
        return int(lib().oracle_count(self.handle, 1))

    def derived(self) -> OracleDerived:
        d = OracleDerived()
        lib().oracle_sim_derived(self.handle, C.byref(d))
        return d

    def step(self, dt: float, pol: OraclePolicy) -> dict:
        st = OracleStats()
        rc = lib().oracle_step(self.handle, dt, C.byref(pol), C.byref(st))
        out = st.as_dict()
        out["status"] = rc
        return out

    def get(self, field: int) -> np.ndarray:
        if field in (F_NBR, F_NBR_B, F_KEYS, F_SORTED):
            out = np.zeros(self.n, np.uint32)
            lib().oracle_get_u32(self.handle, field, _ptr(out))
            return out
        if field == F_CELL_STARTS:
            out = np.zeros(self.derived().total_cells + 1, np.uint32)
            lib().oracle_get_u32(self.handle, field, _ptr(out))
            return out
        count = self.nb if field in _BOUNDARY_FIELDS else self.n
        shape = (count, 3) if field in _VEC_FIELDS else (count,)
        out = np.zeros(shape, np.float64)
        lib().oracle_get(self.handle, field, _ptr(out))
        return out

    def prepare(self, psi: int = 0) -> None:
        """Hash + density/alpha of the current positions (no step)."""
        lib().oracle_prepare(self.handle, psi)

    def apply_A(self, p: np.ndarray, mode_k: int, dt: float, psi: int = 0):
        """(A p, pressure acceleration) of the solver operator after prepare()."""
        pv = np.ascontiguousarray(p, np.float64)
        ap = np.zeros(self.n, np.float64)
        acc = np.zeros((self.n, 3), np.float64)
        lib().oracle_apply_A(self.handle, mode_k, dt, psi, _ptr(pv), _ptr(ap), _ptr(acc))
        return ap, acc

    def source(self, adv: bool, dt: float) -> np.ndarray:
        """One source pass after prepare(): the divergence source rho_dot
        (adv False) or the density source rho* (adv True) of the current state."""
        lib().oracle_source(self.handle, 1 if adv else 0, dt)
        return self.get(F_RHO_ADV if adv else F_RHO_CHANGE)

    def forces(self, dt: float) -> np.ndarray:
        """One force pass after prepare(): v <- v + dt (g + viscosity + cohesion);
        returns the new velocities."""
        lib().oracle_forces(self.handle, dt)
        return self.get(F_V)

    def set(self, field: int, values: np.ndarray) -> None:
        a = np.ascontiguousarray(values, np.float64)
        if lib().oracle_set(self.handle, field, _ptr(a)) != 0:
            raise ValueError(f"field {field} is not settable")


def render(x, view, proj, cam, width, height, radius, point_scale, tpp=0.005, iterations=4, sigma_s=3.0,
           sigma_d=0.05, alternate=False, tint=(0.10, 0.30, 0.50), rough=0.0, zenith=(0.40, 0.55, 0.85),
           horizon=(0.85, 0.90, 0.95), exposure=1.0, f0=0.02, absorption=None, thick_scale=0.0) -> dict:
    """The reference's render passes B-E over positions x (n x 3, draw order):
    rgba (h x w x 4), depth, smoothed, thickness (h x w)."""
    x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
    f32 = lambda a, k: np.ascontiguousarray(a, np.float32).reshape(k)   # noqa: E731
    out = {"rgba": np.zeros((height, width, 4), np.float32), "depth": np.zeros((height, width), np.float32),
           "smoothed": np.zeros((height, width), np.float32), "thickness": np.zeros((height, width), np.float32)}
    args = [f32(view, 16), f32(proj, 16), f32(cam, 3), f32(tint, 3), f32(zenith, 3), f32(horizon, 3)]
    lib().oracle_render(_ptr(x), x.shape[0], _ptr(args[0]), _ptr(args[1]), _ptr(args[2]), width, height, radius,
                        point_scale, tpp, int(iterations), sigma_s, sigma_d, int(bool(alternate)), _ptr(args[3]),
                        rough, _ptr(args[4]), _ptr(args[5]), exposure, f0, tpp if absorption is None else absorption,
                        thick_scale, _ptr(out["rgba"]), _ptr(out["depth"]), _ptr(out["smoothed"]),
                        _ptr(out["thickness"]))
    return out


def render_from(params, x) -> dict:
    """oracle render() with the fields of a dfsph.RenderParams."""
    p = params
    return render(x, list(p.view), list(p.proj), list(p.camera_pos), p.width, p.height, p.particle_radius,
                  p.point_scale, p.thickness_per_particle, p.bilateral_iterations, p.sigma_spatial, p.sigma_depth,
                  p.bilateral_alternate, list(p.water_tint), p.roughness, list(p.sky_zenith), list(p.sky_horizon),
                  p.exposure, p.fresnel_f0, p.absorption, p.thickness_scale)


def render_bilateral(depth, iterations=4, sigma_s=3.0, sigma_d=0.05, alternate=False) -> np.ndarray:
    """The smoothing stage alone on a depth image (h x w)."""
    d = np.ascontiguousarray(depth, np.float32)
    out = np.zeros_like(d)
    lib().oracle_render_bilateral(_ptr(d), d.shape[1], d.shape[0], int(iterations), sigma_s, sigma_d,
                                  int(bool(alternate)), _ptr(out))
    return out


def render_composite(depth, thickness, view, proj, cam, tint=(0.10, 0.30, 0.50), rough=0.0,
                     zenith=(0.40, 0.55, 0.85), horizon=(0.85, 0.90, 0.95), exposure=1.0, f0=0.02,
                     absorption=0.005, thick_scale=0.0) -> np.ndarray:
    """The composite stage alone on a (smoothed) depth and a thickness image."""
    d = np.ascontiguousarray(depth, np.float32)
    t = np.ascontiguousarray(thickness, np.float32)
    f32 = lambda a, k: np.ascontiguousarray(a, np.float32).reshape(k)   # noqa: E731
    v, pr, c, ti, z, hz = f32(view, 16), f32(proj, 16), f32(cam, 3), f32(tint, 3), f32(zenith, 3), f32(horizon, 3)
    out = np.zeros(d.shape + (4,), np.float32)
    lib().oracle_render_composite(_ptr(d), _ptr(t), d.shape[1], d.shape[0], _ptr(v), _ptr(pr), _ptr(c), _ptr(ti),
                                  rough, _ptr(z), _ptr(hz), exposure, f0, absorption, thick_scale, _ptr(out))
    return out
