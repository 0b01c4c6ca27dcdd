"""ctypes mirror of the DFSPH C ABI (include/dfsph.h).

The product is libdfsph.so (HIP kernels for gfx950 behind a C ABI); this
module only marshals numpy arrays across it for tests, smoke checks and the
benchmark.  There is no fallback: if the library is missing or no gfx950 device
is present, calls raise.

Names follow the reference's sph-water vocabulary (particles, density_alpha,
pressure, morton codes, sorted_index, cell_starts; src/main.cpp).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "libdfsph.so")

ABI_VERSION = 1
OK, E_INVALID, E_NOMEM, E_DEVICE, E_NOT_CONVERGED, E_NONFINITE, E_CAPACITY, E_IO = 0, -1, -2, -3, -4, -5, -6, -7
GRID_REFERENCE, GRID_FINE = 0, 1
FLAG_NO_NEIGHBOR_MASKS, FLAG_NO_XCD_REMAP, FLAG_NEIGHBOR_LIST, FLAG_NO_STAGED_WINDOWS = 1, 2, 4, 8
POLICY_REFERENCE, POLICY_CONVERGED = 0, 1
WARM_ZERO, WARM_PREVIOUS, WARM_EXTERNAL, WARM_MLP, WARM_MLP_DENSITY = 0, 1, 2, 3, 4

FIELD_POSITION, FIELD_VELOCITY, FIELD_DENSITY, FIELD_ALPHA_OVER_RHO2 = 0, 1, 2, 3
FIELD_DENSITY_ADV, FIELD_DENSITY_CHANGE, FIELD_PRESSURE_DIV, FIELD_PRESSURE = 4, 5, 6, 7
FIELD_PRESSURE_ACCEL, FIELD_ALPHA, FIELD_NEIGHBOR_COUNT, FIELD_BOUNDARY_NEIGHBOR_COUNT = 8, 9, 10, 11
FIELD_MORTON_KEY, FIELD_SORTED_INDEX, FIELD_CELL_STARTS = 12, 13, 14
FIELD_BOUNDARY_POSITION, FIELD_BOUNDARY_VOLUME, FIELD_ID = 15, 16, 21

_VEC = {FIELD_POSITION, FIELD_VELOCITY, FIELD_PRESSURE_ACCEL, FIELD_BOUNDARY_POSITION}
_U32 = {FIELD_NEIGHBOR_COUNT, FIELD_BOUNDARY_NEIGHBOR_COUNT, FIELD_MORTON_KEY,
        FIELD_SORTED_INDEX, FIELD_CELL_STARTS, FIELD_ID}

# Every symbol include/dfsph.h declares.
EXPORTED = [
    "dfsph_create", "dfsph_destroy", "dfsph_init_scene", "dfsph_step", "dfsph_field_size",
    "dfsph_export", "dfsph_set_warm_start", "dfsph_last_error", "dfsph_default_policy",
    "dfsph_derive", "dfsph_boundary_samples", "dfsph_scene_block", "dfsph_profile_enable",
    "dfsph_profile_read", "dfsph_mlp_init", "dfsph_mlp_set_params", "dfsph_mlp_get_params",
    "dfsph_dataset_start", "dfsph_dataset_stop", "dfsph_dataset_export", "dfsph_dataset_import",
    "dfsph_mlp_train", "dfsph_mlp_eval", "dfsph_mlp_grad", "dfsph_mlp_grad_noalloc",
    "dfsph_batch_create", "dfsph_batch_destroy", "dfsph_batch_info", "dfsph_batch_step",
    "dfsph_batch_field_size", "dfsph_batch_export", "dfsph_autotune", "dfsph_batch_autotune",
    "dfsph_capture_save", "dfsph_capture_load", "dfsph_capture_read_info", "dfsph_capture_find_latest",
    "dfsph_preset_get", "dfsph_preset_fill", "dfsph_emit", "dfsph_trace_mark", "dfsph_batch_trace_mark",
    "dfsph_abc_open", "dfsph_abc_write", "dfsph_abc_write_ctx", "dfsph_abc_close",
    "dfsph_render_defaults", "dfsph_render",
]
PRESET_COUNT, MAX_EMITTERS = 4, 8
DIV_START_DEFAULT, DIV_START_ZERO, DIV_START_PREVIOUS, DIV_START_MLP = 0, 1, 2, 3
BATCH_NO_GRAPH, BATCH_NO_XCD_MAP = 1, 2
MLP_PARAMS, MLP_FEATURES, MLP_OUTPUTS = 9538, 16, 2
MLP_GRAD_NOALLOC_MAX_ROWS = 4096


class Params(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32), ("capacity", C.c_uint32), ("particle_radius", C.c_float),
        ("density0", C.c_float), ("gravity", C.c_float * 3), ("viscosity", C.c_float),
        ("cohesion", C.c_float), ("vorticity", C.c_float), ("domain_min", C.c_float * 3),
        ("domain_max", C.c_float * 3), ("grid", C.c_int32), ("device", C.c_int32),
        ("block_size", C.c_uint32), ("flags", C.c_uint32), ("fine_cell_scale", C.c_float),
        ("reserved", C.c_uint32 * 3),
    ]


class Scene(C.Structure):
    _fields_ = [
        ("count", C.c_uint32), ("positions", C.c_void_p), ("velocities", C.c_void_p),
        ("ids", C.c_void_p), ("boundary", C.c_int32), ("boundary_cap", C.c_uint32),
        ("flags", C.c_void_p),
    ]


class Preset(C.Structure):
    _fields_ = [
        ("name", C.c_char * 32), ("brick_min", C.c_float * 3), ("brick_max", C.c_float * 3),
        ("add_droplet", C.c_int32), ("droplet_center", C.c_float * 3), ("droplet_radius", C.c_float),
        ("droplet_velocity", C.c_float * 3), ("source_shape", C.c_int32), ("source_pos", C.c_float * 3),
        ("source_size", C.c_float * 3), ("source_velocity", C.c_float * 3), ("source_rate", C.c_float),
        ("domain_min", C.c_float * 3), ("domain_max", C.c_float * 3), ("viscosity", C.c_float),
        ("cohesion", C.c_float), ("vorticity", C.c_float), ("roughness", C.c_float),
    ]


class Emitter(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("radius", C.c_float), ("velocity", C.c_float * 3), ("rate", C.c_float)]


class RenderParams(C.Structure):
    _fields_ = [
        ("view", C.c_float * 16), ("proj", C.c_float * 16), ("camera_pos", C.c_float * 3),
        ("width", C.c_uint32), ("height", C.c_uint32), ("particle_radius", C.c_float),
        ("point_scale", C.c_float), ("thickness_per_particle", C.c_float),
        ("bilateral_iterations", C.c_int32), ("sigma_spatial", C.c_float), ("sigma_depth", C.c_float),
        ("bilateral_alternate", C.c_int32), ("water_tint", C.c_float * 3), ("roughness", C.c_float),
        ("sky_zenith", C.c_float * 3), ("sky_horizon", C.c_float * 3), ("exposure", C.c_float),
        ("fresnel_f0", C.c_float), ("absorption", C.c_float), ("thickness_scale", C.c_float),
        ("reserved", C.c_uint32 * 8),
    ]


class CaptureInfo(C.Structure):
    _fields_ = [
        ("iteration", C.c_uint64), ("frame", C.c_uint32), ("particle_count", C.c_uint32),
        ("tier_index", C.c_int32), ("preset_index", C.c_int32), ("substeps", C.c_int32),
        ("reserved0", C.c_int32), ("camera_json", C.c_char_p), ("reserved", C.c_int32 * 4),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k in ("iteration", "frame", "particle_count", "tier_index",
                                              "preset_index", "substeps")}


class Policy(C.Structure):
    _fields_ = [
        ("mode", C.c_int32), ("warm_start", C.c_int32), ("div_enabled", C.c_int32),
        ("min_iter_div", C.c_int32), ("max_iter_div", C.c_int32),
        ("min_iter_density", C.c_int32), ("max_iter_density", C.c_int32),
        ("tol_div", C.c_float), ("tol_density", C.c_float), ("cfl", C.c_float),
        ("dt_max", C.c_float), ("boundary_psi", C.c_int32), ("strict", C.c_int32),
        ("residual_bound", C.c_float), ("div_start", C.c_int32), ("reserved", C.c_int32 * 1),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("iters_div", C.c_int32), ("iters_density", C.c_int32),
        ("converged_div", C.c_int32), ("converged_density", C.c_int32),
        ("err_div", C.c_double), ("err_density", C.c_double), ("max_speed", C.c_double),
        ("nonfinite", C.c_uint64), ("gpu_ms", C.c_double), ("launches", C.c_uint32),
        ("host_syncs", C.c_uint32), ("warm_fallback", C.c_int32),
        ("max_residual_div", C.c_float), ("max_residual_density", C.c_float), ("reserved", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class Derived(C.Structure):
    _fields_ = [
        ("h", C.c_float), ("cell", C.c_float), ("mass", C.c_float), ("knorm", C.c_float),
        ("gnorm", C.c_float), ("relax", C.c_float), ("axes", C.c_uint32 * 3),
        ("max_axis", C.c_uint32), ("total_cells", C.c_uint32), ("fine_cell", C.c_float),
        ("fine_axes", C.c_uint32 * 3), ("r2_q1", C.c_float), ("r2_eps_gt", C.c_float),
        ("r2_eps_lt", C.c_float),
    ]


class KernelTime(C.Structure):
    _fields_ = [
        ("name", C.c_char * 32), ("launches", C.c_uint64), ("total_ms", C.c_double),
        ("min_ms", C.c_double), ("max_ms", C.c_double),
    ]


class TrainConfig(C.Structure):
    _fields_ = [
        ("epochs", C.c_uint32), ("batch", C.c_uint32), ("lr", C.c_float), ("beta1", C.c_float),
        ("beta2", C.c_float), ("eps", C.c_float), ("seed", C.c_uint64), ("val_fraction", C.c_float),
        ("expectile_density", C.c_float), ("reserved", C.c_uint32 * 2),
    ]


class TrainStats(C.Structure):
    _fields_ = [
        ("train_loss", C.c_double), ("val_loss", C.c_double), ("steps", C.c_uint32),
        ("train_rows", C.c_uint32), ("val_rows", C.c_uint32), ("reserved", C.c_uint32),
        ("ms", C.c_double),
    ]


class BatchConfig(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32), ("device", C.c_int32), ("workgroups", C.c_uint32),
        ("flags", C.c_uint32), ("spin_limit", C.c_uint32), ("reserved", C.c_uint32 * 3),
    ]


class BatchInfo(C.Structure):
    _fields_ = [
        ("scenes", C.c_uint32), ("particles", C.c_uint32), ("boundary_particles", C.c_uint32),
        ("tiles", C.c_uint32), ("cells", C.c_uint32), ("workgroups", C.c_uint32),
        ("max_workgroups", C.c_uint32), ("reserved", C.c_uint32),
    ]


class TuneResult(C.Structure):
    _fields_ = [
        ("block_size", C.c_uint32), ("flags", C.c_uint32), ("workgroups", C.c_uint32),
        ("batch_flags", C.c_uint32), ("ms_per_step", C.c_double), ("worst_ms_per_step", C.c_double),
        ("candidates", C.c_uint32), ("cached", C.c_int32), ("fine_cell_scale", C.c_float),
        ("reserved", C.c_uint32),
    ]


class DfsphError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"dfsph error {code}: {msg}")
        self.code = code


def build() -> str:
    subprocess.run(["make", "-s", "-C", PKG, "-j8"], check=True)
    return LIB_PATH


_lib = None


def lib() -> C.CDLL:
    """Load libdfsph.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built; run `make -C gpu-sims_amd`")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.dfsph_last_error.restype = C.c_char_p
        L.dfsph_create.argtypes = [C.POINTER(Params), C.POINTER(C.c_void_p)]
        L.dfsph_destroy.argtypes = [P]
        L.dfsph_init_scene.argtypes = [P, C.POINTER(Scene)]
        L.dfsph_step.argtypes = [P, C.c_float, C.c_int32, C.POINTER(Policy), C.POINTER(StepStats)]
        L.dfsph_field_size.argtypes = [P, C.c_int32, C.POINTER(C.c_size_t)]
        L.dfsph_export.argtypes = [P, C.c_int32, P, C.c_size_t]
        L.dfsph_set_warm_start.argtypes = [P, P, P]
        L.dfsph_default_policy.argtypes = [C.c_int32, C.POINTER(Policy)]
        L.dfsph_derive.argtypes = [C.POINTER(Params), C.POINTER(Derived)]
        L.dfsph_boundary_samples.argtypes = [P, P, C.c_float, C.c_uint32, P, C.c_uint32,
                                             C.POINTER(C.c_uint32)]
        L.dfsph_scene_block.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, P,
                                        C.c_uint64, C.c_float, C.c_uint32, P]
        L.dfsph_profile_enable.argtypes = [P, C.c_int32]
        L.dfsph_profile_read.argtypes = [P, P, C.c_uint32, C.POINTER(C.c_uint32)]
        L.dfsph_trace_mark.argtypes = [P, C.c_uint32]
        L.dfsph_abc_open.argtypes = [C.c_char_p, C.c_double, C.POINTER(C.c_void_p)]
        L.dfsph_abc_write.argtypes = [P, P, P, P, C.c_uint32]
        L.dfsph_abc_write_ctx.argtypes = [P, P]
        L.dfsph_abc_close.argtypes = [P]
        L.dfsph_render_defaults.argtypes = [P, P, C.c_float, C.c_uint32, C.c_uint32, C.POINTER(RenderParams)]
        L.dfsph_render.argtypes = [P, C.POINTER(RenderParams), P, P, P, P]
        L.dfsph_batch_trace_mark.argtypes = [P, C.c_uint32]
        L.dfsph_mlp_init.argtypes = [P, C.c_uint64]
        L.dfsph_mlp_set_params.argtypes = [P, P, C.c_uint32]
        L.dfsph_mlp_get_params.argtypes = [P, P, C.c_uint32]
        L.dfsph_dataset_start.argtypes = [P, C.c_uint32, C.c_uint32]
        L.dfsph_dataset_stop.argtypes = [P, C.POINTER(C.c_uint32)]
        L.dfsph_dataset_export.argtypes = [P, P, P, C.c_uint32]
        L.dfsph_dataset_import.argtypes = [P, P, P, C.c_uint32]
        L.dfsph_mlp_train.argtypes = [P, C.POINTER(TrainConfig), C.POINTER(TrainStats)]
        L.dfsph_mlp_eval.argtypes = [P, P, C.c_uint32, P]
        L.dfsph_mlp_grad.argtypes = [P, P, P, C.c_uint32, P, C.POINTER(C.c_double)]
        L.dfsph_mlp_grad_noalloc.argtypes = [P, P, P, C.c_uint32, P, C.POINTER(C.c_double)]
        L.dfsph_batch_create.argtypes = [C.POINTER(BatchConfig), P, P, C.c_uint32, C.POINTER(C.c_void_p)]
        L.dfsph_batch_destroy.argtypes = [P]
        L.dfsph_batch_info.argtypes = [P, C.POINTER(BatchInfo)]
        L.dfsph_batch_step.argtypes = [P, C.c_float, C.POINTER(Policy), P, C.POINTER(StepStats)]
        L.dfsph_batch_field_size.argtypes = [P, C.c_uint32, C.c_int32, C.POINTER(C.c_size_t)]
        L.dfsph_batch_export.argtypes = [P, C.c_uint32, C.c_int32, P, C.c_size_t]
        L.dfsph_capture_save.argtypes = [P, C.c_char_p, C.c_uint32, C.POINTER(CaptureInfo), C.c_char_p,
                                         C.c_size_t]
        L.dfsph_capture_load.argtypes = [P, C.c_char_p, C.c_int32, C.c_uint32, C.POINTER(CaptureInfo)]
        L.dfsph_capture_read_info.argtypes = [C.c_char_p, C.POINTER(CaptureInfo)]
        L.dfsph_capture_find_latest.argtypes = [C.c_char_p, C.c_char_p, C.c_size_t]
        L.dfsph_preset_get.argtypes = [C.c_int32, C.POINTER(Preset)]
        L.dfsph_preset_fill.argtypes = [C.POINTER(Preset), C.c_float, C.c_uint32, P, P, P,
                                        C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.dfsph_emit.argtypes = [P, P, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32)]
        L.dfsph_autotune.argtypes = [C.POINTER(Params), C.POINTER(Scene), C.POINTER(Policy), C.c_uint32,
                                     C.c_char_p, C.POINTER(TuneResult)]
        L.dfsph_batch_autotune.argtypes = [C.POINTER(BatchConfig), P, P, C.c_uint32, C.POINTER(Policy),
                                           C.c_uint32, C.c_char_p, C.POINTER(TuneResult)]
        for name in EXPORTED:
            if name != "dfsph_last_error":
                getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != OK:
        raise DfsphError(rc, lib().dfsph_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def make_params(r, dmin, dmax, capacity, rho0=1000.0, gravity=(0.0, -9.81, 0.0), viscosity=0.01,
                cohesion=0.0, vorticity=0.0, grid=GRID_FINE, device=0, block_size=0,
                flags=0, fine_cell_scale=0.0) -> Params:
    p = Params()
    p.abi_version = ABI_VERSION
    p.capacity = capacity
    p.particle_radius = r
    p.density0 = rho0
    p.gravity[:] = list(gravity)
    p.viscosity, p.cohesion, p.vorticity = viscosity, cohesion, vorticity
    p.domain_min[:] = list(dmin)
    p.domain_max[:] = list(dmax)
    p.grid, p.device, p.block_size, p.flags = grid, device, block_size, flags
    p.fine_cell_scale = fine_cell_scale
    return p


def default_policy(mode=POLICY_REFERENCE, **kw) -> Policy:
    p = Policy()
    _check(lib().dfsph_default_policy(mode, C.byref(p)))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def derive(params: Params) -> Derived:
    d = Derived()
    _check(lib().dfsph_derive(C.byref(params), C.byref(d)))
    return d


def boundary_samples(dmin, dmax, r, cap=0) -> np.ndarray:
    lo = np.asarray(dmin, np.float32)
    hi = np.asarray(dmax, np.float32)
    cnt = C.c_uint32()
    _check(lib().dfsph_boundary_samples(_ptr(lo), _ptr(hi), r, cap, None, 0, C.byref(cnt)))
    out = np.zeros((cnt.value, 3), np.float32)
    _check(lib().dfsph_boundary_samples(_ptr(lo), _ptr(hi), r, cap, _ptr(out), cnt.value, C.byref(cnt)))
    return out


def scene_block(nx, ny, nz, r, origin, seed, jitter=0.1, id0=0) -> np.ndarray:
    o = np.asarray(origin, np.float32)
    out = np.zeros((nx * ny * nz, 3), np.float32)
    _check(lib().dfsph_scene_block(nx, ny, nz, r, _ptr(o), seed, jitter, id0, _ptr(out)))
    return out


class Sim:
    """One DFSPH context on one GPU (mirrors the reference's init/step/export surface)."""

    def __init__(self, params: Params):
        self.params = params
        h = C.c_void_p()
        _check(lib().dfsph_create(C.byref(params), C.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            lib().dfsph_destroy(self.handle)
            self.handle = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def init_scene(self, x: np.ndarray, v: np.ndarray | None = None, ids: np.ndarray | None = None,
                   boundary=True, boundary_cap=0, flags: np.ndarray | None = None) -> None:
        x = np.ascontiguousarray(x, np.float32)
        v = None if v is None else np.ascontiguousarray(v, np.float32)
        ids = None if ids is None else np.ascontiguousarray(ids, np.uint32)
        flags = None if flags is None else np.ascontiguousarray(flags, np.uint32)
        s = Scene()
        s.count = x.shape[0]
        s.positions, s.velocities, s.ids, s.flags = _ptr(x), _ptr(v), _ptr(ids), _ptr(flags)
        s.boundary = 1 if boundary else 0
        s.boundary_cap = boundary_cap
        _check(lib().dfsph_init_scene(self.handle, C.byref(s)))

    def emit(self, emitters, n_new: int) -> int:
        """Append up to n_new particles from emitters given as (pos xyz, radius,
        velocity xyz, rate) rows; returns how many were added."""
        em = np.ascontiguousarray(emitters, np.float32).reshape(-1, 8)
        arr = (Emitter * em.shape[0])()
        for k, row in enumerate(em):
            arr[k].pos[:] = row[0:3]
            arr[k].radius = row[3]
            arr[k].velocity[:] = row[4:7]
            arr[k].rate = row[7]
        out = C.c_uint32()
        _check(lib().dfsph_emit(self.handle, C.cast(arr, C.c_void_p), em.shape[0], n_new, C.byref(out)))
        return out.value

    def render(self, params: "RenderParams", rgba=True, depth=True, smoothed=True, thickness=True) -> dict:
        """Screen-space render passes of the current particles: a dict of the
        requested images (rgba height x width x 4, the others height x width)."""
        h, w = params.height, params.width
        out = {}
        for name, want, shape in (("rgba", rgba, (h, w, 4)), ("depth", depth, (h, w)),
                                  ("smoothed", smoothed, (h, w)), ("thickness", thickness, (h, w))):
            out[name] = np.zeros(shape, np.float32) if want else None
        _check(lib().dfsph_render(self.handle, C.byref(params), _ptr(out["rgba"]), _ptr(out["depth"]),
                                  _ptr(out["smoothed"]), _ptr(out["thickness"])))
        return {k: v for k, v in out.items() if v is not None}

    def capture_save(self, root: str, frame: int, iteration: int = 0, tier_index: int = 0,
                     preset_index: int = 0, substeps: int = 1, camera_json: str | None = None) -> str:
        """Write <root>/capture_NNNN in the sphWater capture format; returns its path."""
        info = CaptureInfo()
        info.iteration, info.tier_index, info.preset_index, info.substeps = iteration, tier_index, \
            preset_index, substeps
        info.camera_json = None if camera_json is None else camera_json.encode()
        out = C.create_string_buffer(4096)
        _check(lib().dfsph_capture_save(self.handle, str(root).encode(), frame, C.byref(info), out, 4096))
        return out.value.decode()

    def capture_load(self, directory: str, boundary=True, boundary_cap=0) -> dict:
        """Re-initialise the scene from a capture directory; returns its meta."""
        info = CaptureInfo()
        _check(lib().dfsph_capture_load(self.handle, str(directory).encode(), 1 if boundary else 0,
                                        boundary_cap, C.byref(info)))
        return info.as_dict()

    def step(self, policy: Policy, dt: float = 0.0, substeps: int = 1, check=True) -> dict:
        st = StepStats()
        rc = lib().dfsph_step(self.handle, dt, substeps, C.byref(policy), C.byref(st))
        out = st.as_dict()
        out["status"] = rc
        if check:
            _check(rc)
        return out

    def set_warm_start(self, p_div=None, p_density=None) -> None:
        a = None if p_div is None else np.ascontiguousarray(p_div, np.float32)
        b = None if p_density is None else np.ascontiguousarray(p_density, np.float32)
        _check(lib().dfsph_set_warm_start(self.handle, _ptr(a), _ptr(b)))

    def profile(self, enable: bool = True) -> None:
        """Enable (and reset) per-kernel HIP-event timing on the context stream."""
        _check(lib().dfsph_profile_enable(self.handle, int(enable)))

    def trace_mark(self, tag: int = 0) -> None:
        """One empty k_trace_mark launch on the context stream (a region marker)."""
        _check(lib().dfsph_trace_mark(self.handle, tag))

    def kernel_times(self) -> dict:
        cnt = C.c_uint32()
        _check(lib().dfsph_profile_read(self.handle, None, 0, C.byref(cnt)))
        arr = (KernelTime * max(cnt.value, 1))()
        _check(lib().dfsph_profile_read(self.handle, C.cast(arr, C.c_void_p), cnt.value, C.byref(cnt)))
        return {arr[i].name.decode(): {"launches": arr[i].launches, "total_ms": arr[i].total_ms,
                                       "avg_ms": arr[i].total_ms / max(arr[i].launches, 1),
                                       "min_ms": arr[i].min_ms, "max_ms": arr[i].max_ms}
                for i in range(cnt.value)}

    # ---- learned warm start ----
    def mlp_init(self, seed: int) -> None:
        _check(lib().dfsph_mlp_init(self.handle, seed))

    def mlp_set_params(self, theta: np.ndarray) -> None:
        t = np.ascontiguousarray(theta, np.float32)
        _check(lib().dfsph_mlp_set_params(self.handle, _ptr(t), t.size))

    def mlp_get_params(self) -> np.ndarray:
        t = np.zeros(MLP_PARAMS, np.float32)
        _check(lib().dfsph_mlp_get_params(self.handle, _ptr(t), t.size))
        return t

    def dataset_start(self, capacity_rows: int, stride: int) -> None:
        _check(lib().dfsph_dataset_start(self.handle, capacity_rows, stride))

    def dataset_stop(self) -> int:
        rows = C.c_uint32()
        _check(lib().dfsph_dataset_stop(self.handle, C.byref(rows)))
        return rows.value

    def dataset_export(self, rows: int):
        x = np.zeros((rows, MLP_FEATURES), np.float32)
        t = np.zeros((rows, MLP_OUTPUTS), np.float32)
        _check(lib().dfsph_dataset_export(self.handle, _ptr(x), _ptr(t), rows))
        return x, t

    def dataset_import(self, x: np.ndarray, t: np.ndarray) -> None:
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        _check(lib().dfsph_dataset_import(self.handle, _ptr(x), _ptr(t), x.shape[0]))

    def mlp_train(self, epochs=10, batch=4096, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, seed=1,
                  val_fraction=0.2, expectile_density=0.0) -> dict:
        cfg = TrainConfig()
        cfg.epochs, cfg.batch, cfg.lr, cfg.beta1, cfg.beta2, cfg.eps = epochs, batch, lr, beta1, beta2, eps
        cfg.seed, cfg.val_fraction, cfg.expectile_density = seed, val_fraction, expectile_density
        st = TrainStats()
        _check(lib().dfsph_mlp_train(self.handle, C.byref(cfg), C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_ if k != "reserved"}

    def mlp_eval(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros((x.shape[0], MLP_OUTPUTS), np.float32)
        _check(lib().dfsph_mlp_eval(self.handle, _ptr(x), x.shape[0], _ptr(out)))
        return out

    def mlp_grad(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_double()
        _check(lib().dfsph_mlp_grad(self.handle, _ptr(x), _ptr(t), x.shape[0], _ptr(g), C.byref(loss)))
        return loss.value, g

    def mlp_grad_noalloc(self, x: np.ndarray, t: np.ndarray):
        """dfsph_mlp_grad_noalloc: the same as mlp_grad without any allocation
        (rows up to MLP_GRAD_NOALLOC_MAX_ROWS)."""
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_double()
        _check(lib().dfsph_mlp_grad_noalloc(self.handle, _ptr(x), _ptr(t), x.shape[0], _ptr(g),
                                            C.byref(loss)))
        return loss.value, g

    def field_size(self, field: int) -> int:
        n = C.c_size_t()
        _check(lib().dfsph_field_size(self.handle, field, C.byref(n)))
        return n.value

    def export(self, field: int) -> np.ndarray:
        nbytes = self.field_size(field)
        dtype = np.uint32 if field in _U32 else np.float32
        out = np.zeros(nbytes // 4, dtype)
        _check(lib().dfsph_export(self.handle, field, _ptr(out), nbytes))
        return out.reshape(-1, 3) if field in _VEC else out


def preset(index: int) -> Preset:
    p = Preset()
    _check(lib().dfsph_preset_get(index, C.byref(p)))
    return p


def preset_fill(p: Preset, r: float, capacity: int):
    """Initial fill of a preset: (positions, velocities, flags, brick_count)."""
    n, nb = C.c_uint32(), C.c_uint32()
    _check(lib().dfsph_preset_fill(C.byref(p), C.c_float(r), capacity, None, None, None, C.byref(n), C.byref(nb)))
    x = np.zeros((n.value, 3), np.float32)
    v = np.zeros((n.value, 3), np.float32)
    f = np.zeros(n.value, np.uint32)
    _check(lib().dfsph_preset_fill(C.byref(p), C.c_float(r), capacity, _ptr(x), _ptr(v), _ptr(f), C.byref(n),
                                   C.byref(nb)))
    return x, v, f, nb.value


def render_defaults(eye, target, fov_deg: float, width: int, height: int, **kw) -> RenderParams:
    """The reference's render settings for a camera at eye looking at target."""
    p = RenderParams()
    e = np.ascontiguousarray(eye, np.float32)
    t = np.ascontiguousarray(target, np.float32)
    _check(lib().dfsph_render_defaults(_ptr(e), _ptr(t), float(fov_deg), width, height, C.byref(p)))
    for k, v in kw.items():
        if isinstance(getattr(p, k), C.Array):
            getattr(p, k)[:] = list(v)
        else:
            setattr(p, k, v)
    return p


class AbcWriter:
    """Alembic particle export (the reference's abc::ParticleWriter): one
    OPoints object per archive, a frame per write, complete after close()."""

    def __init__(self, path: str, fps: float = 30.0):
        h = C.c_void_p()
        _check(lib().dfsph_abc_open(os.fsencode(path), float(fps), C.byref(h)))
        self.handle = h

    def write(self, positions, velocities=None, ids=None) -> None:
        x = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        v = None if velocities is None else np.ascontiguousarray(velocities, np.float32).reshape(-1, 3)
        i = None if ids is None else np.ascontiguousarray(ids, np.uint64)
        if v is not None and v.shape != x.shape or i is not None and i.shape != (x.shape[0],):
            raise ValueError("velocities must match positions and ids must hold one entry per particle")
        _check(lib().dfsph_abc_write(self.handle, _ptr(x), _ptr(v), _ptr(i), x.shape[0]))

    def write_sim(self, sim: "Sim") -> None:
        _check(lib().dfsph_abc_write_ctx(self.handle, sim.handle))

    def close(self) -> None:
        if self.handle:
            h, self.handle = self.handle, None
            _check(lib().dfsph_abc_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def capture_read_info(directory: str) -> dict:
    """Check a capture directory as the reference's loader does (no GPU needed)."""
    info = CaptureInfo()
    _check(lib().dfsph_capture_read_info(str(directory).encode(), C.byref(info)))
    return info.as_dict()


def capture_find_latest(root: str) -> str:
    out = C.create_string_buffer(4096)
    _check(lib().dfsph_capture_find_latest(str(root).encode(), out, 4096))
    return out.value.decode()


def batch_config(workgroups=0, flags=0, device=0, spin_limit=0) -> BatchConfig:
    c = BatchConfig()
    c.abi_version = ABI_VERSION
    c.device = device
    c.workgroups = workgroups
    c.flags = flags
    c.spin_limit = spin_limit
    return c


def _scene_arrays(params_list, scenes):
    """Keep-alive tuple of ctypes arrays (Params[], Scene[]) for scenes given as
    (x[, v[, ids]]) numpy arrays."""
    n = len(scenes)
    P = (Params * n)(*params_list)
    S = (Scene * n)()
    keep = []
    for k, sc in enumerate(scenes):
        x = np.ascontiguousarray(sc[0], np.float32)
        v = np.ascontiguousarray(sc[1], np.float32) if len(sc) > 1 and sc[1] is not None else None
        ids = np.ascontiguousarray(sc[2], np.uint32) if len(sc) > 2 and sc[2] is not None else None
        keep += [x, v, ids]
        S[k].count = x.shape[0]
        S[k].positions, S[k].velocities, S[k].ids = _ptr(x), _ptr(v), _ptr(ids)
        S[k].boundary = 1
        S[k].boundary_cap = 0
    return P, S, keep


class Batch:
    """Many scenes advanced by one persistent launch per substep (dfsph_batch_*)."""

    def __init__(self, params_list, scenes, workgroups=0, flags=0, device=0, spin_limit=0):
        self.n_scenes = len(scenes)
        P, S, keep = _scene_arrays(params_list, scenes)
        self.counts = [int(np.asarray(sc[0]).shape[0]) for sc in scenes]
        h = C.c_void_p()
        _check(lib().dfsph_batch_create(C.byref(batch_config(workgroups, flags, device, spin_limit)), P, S,
                                        self.n_scenes, C.byref(h)))
        self.handle = h
        del keep

    def close(self):
        if self.handle:
            lib().dfsph_batch_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def info(self) -> dict:
        i = BatchInfo()
        _check(lib().dfsph_batch_info(self.handle, C.byref(i)))
        return {k: getattr(i, k) for k, _ in i._fields_ if k != "reserved"}

    def trace_mark(self, tag: int = 0) -> None:
        """One empty k_trace_mark launch on the batch's stream (a region marker)."""
        _check(lib().dfsph_batch_trace_mark(self.handle, tag))

    def step(self, policy: Policy, dt: float = 0.0, check=True):
        per = (StepStats * self.n_scenes)()
        tot = StepStats()
        rc = lib().dfsph_batch_step(self.handle, dt, C.byref(policy), per, C.byref(tot))
        if check:
            _check(rc)
        t = tot.as_dict()
        t["status"] = rc
        return [per[k].as_dict() for k in range(self.n_scenes)], t

    def field_size(self, scene: int, field: int) -> int:
        n = C.c_size_t()
        _check(lib().dfsph_batch_field_size(self.handle, scene, field, C.byref(n)))
        return n.value

    def export(self, scene: int, field: int) -> np.ndarray:
        nbytes = self.field_size(scene, field)
        dt = np.uint32 if field in _U32 else np.float32
        out = np.empty(nbytes // 4, dt)
        _check(lib().dfsph_batch_export(self.handle, scene, field, out.ctypes.data_as(C.c_void_p), nbytes))
        return out.reshape(-1, 3) if field in _VEC else out


def autotune(params: Params, x: np.ndarray, policy: Policy, steps=5, cache_path=None) -> dict:
    x = np.ascontiguousarray(x, np.float32)
    sc = Scene()
    sc.count = x.shape[0]
    sc.positions = _ptr(x)
    sc.boundary = 1
    r = TuneResult()
    _check(lib().dfsph_autotune(C.byref(params), C.byref(sc), C.byref(policy), steps,
                                cache_path.encode() if cache_path else None, C.byref(r)))
    return {k: getattr(r, k) for k, _ in r._fields_ if k != "reserved"}


def batch_autotune(params_list, scenes, policy: Policy, steps=5, cache_path=None, flags=0, device=0) -> dict:
    P, S, keep = _scene_arrays(params_list, scenes)
    r = TuneResult()
    _check(lib().dfsph_batch_autotune(C.byref(batch_config(0, flags, device)), P, S, len(scenes),
                                      C.byref(policy), steps, cache_path.encode() if cache_path else None,
                                      C.byref(r)))
    del keep
    return {k: getattr(r, k) for k, _ in r._fields_ if k != "reserved"}
