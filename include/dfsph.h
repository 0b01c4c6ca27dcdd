/*
 * dfsph.h -- C ABI of the MI355X (gfx950) DFSPH water substep.
 *
 * Drop-in for the sph-water hot path of StevenFAU/GPU-Sims
 * (particle-fluids/sph-water).  The reference has no plugin or FFI surface:
 * its init, step and export live inline in main() (src/main.cpp).  Each entry
 * point below names the reference code it replaces.  Plain C types only:
 * pointers, sizes, fixed-width integers and floats; no exception crosses it;
 * every call returns a dfsph_status and dfsph_last_error() holds the message.
 *
 * Threading: a context owns one HIP stream and is not thread-safe; separate
 * contexts may be driven from separate host threads.  Caller pointers are
 * borrowed for the duration of the call only (the library owns device memory).
 */
#ifndef DFSPH_H
#define DFSPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DFSPH_ABI_VERSION 1u

/* Status codes (the reference throws in setup, main.cpp:1279, and logs and
 * returns on I/O, main.cpp:2490-2527; here every failure is a code). */
typedef enum dfsph_status {
    DFSPH_OK = 0,
    DFSPH_E_INVALID = -1,        /* bad argument, size or state */
    DFSPH_E_NOMEM = -2,          /* device or host allocation failed */
    DFSPH_E_DEVICE = -3,         /* HIP runtime error or no usable gfx950 device */
    DFSPH_E_NOT_CONVERGED = -4,  /* a solver hit max_iter above tolerance (strict policy) */
    DFSPH_E_NONFINITE = -5,      /* NaN/Inf in position or velocity after the step */
    DFSPH_E_CAPACITY = -6,       /* particle or boundary count above capacity */
    DFSPH_E_IO = -7              /* capture file missing, short or mismatched */
} dfsph_status;

/* Neighbour grid used by the step.
 *   REFERENCE: 30-bit Morton keys over cells of 2h with power-of-two axes and a
 *              maxAxis^3 table (shaders/morton_code.comp.glsl:24-50,
 *              main.cpp:1994-2047); neighbours walked as 27 cells.
 *   FINE:      linear (x fastest) cell index over cells of ~h; neighbours
 *              walked as 9 contiguous x-runs of 3 cells.  Same neighbour sets
 *              (any cell >= h finds every pair with r < h), fewer candidates. */
typedef enum dfsph_grid { DFSPH_GRID_REFERENCE = 0, DFSPH_GRID_FINE = 1 } dfsph_grid;

typedef struct dfsph_ctx dfsph_ctx;

/* Simulation constants.  Replaces the Runtime struct and constexprs of
 * main.cpp:104-156, :246-293 and the UBO packing of main.cpp:2021-2105. */
typedef struct dfsph_params {
    uint32_t abi_version;     /* must be DFSPH_ABI_VERSION */
    uint32_t capacity;        /* maximum fluid particles */
    float particle_radius;    /* r; h = 4r (main.cpp:269), Morton cell 8r (main.cpp:270) */
    float density0;           /* rest density, 1000 (main.cpp:128) */
    float gravity[3];         /* (0,-9.81,0) (main.cpp:2092-2095) */
    float viscosity;          /* main.cpp:2085 */
    float cohesion;           /* main.cpp:2086 */
    float vorticity;          /* carried; the reference's term is a no-op
                                 (shaders/integrate_forces.comp.glsl:137-139) */
    float domain_min[3];      /* fluid AABB (main.cpp:303-304) */
    float domain_max[3];
    int32_t grid;             /* dfsph_grid */
    int32_t device;           /* HIP device ordinal */
    uint32_t block_size;      /* threads per workgroup in the sweeps; 0 = default */
    uint32_t flags;           /* DFSPH_FLAG_* tuning switches; 0 = defaults */
    float fine_cell_scale;    /* FINE grid cell edge in units of h: 0 = 1 + 2^-10 (default),
                                 else in [1 + 2^-10, 4].  Larger cells hold more candidates
                                 per x-run; neighbour sets and results do not depend on it
                                 beyond summation order. */
    uint32_t reserved[3];     /* zero */
} dfsph_params;

/* Tuning switches (results are identical either way; only speed differs). */
#define DFSPH_FLAG_NO_NEIGHBOR_MASKS 1u  /* FINE grid: re-test all candidates in every sweep */
#define DFSPH_FLAG_NO_XCD_REMAP      2u  /* plain workgroup -> tile order */
#define DFSPH_FLAG_NEIGHBOR_LIST     4u  /* FINE grid: per-particle index lists instead of masks */
#define DFSPH_FLAG_NO_STAGED_WINDOWS 8u  /* FINE grid with masks: the solver-phase sweeps gather
                                            neighbours from memory instead of staging each wave's
                                            run windows in LDS */

/* Initial scene.  Replaces seedInitialFill (main.cpp:2266-2294) and
 * setupBoundarySpatialHash (main.cpp:2306-2372): the caller supplies the fluid
 * particles; the library samples the Akinci boundary on the domain AABB
 * (main.cpp:1332-1431), sorts it and computes V_b on the device
 * (shaders/compute_boundary_volume.comp.glsl:55-92). */
typedef struct dfsph_scene {
    uint32_t count;
    const float* positions;   /* count x 3, xyz interleaved */
    const float* velocities;  /* count x 3, or NULL for rest */
    const uint32_t* ids;      /* original indices, or NULL for 0..count-1 */
    int32_t boundary;         /* 1: Akinci samples on the domain AABB; 0: none */
    uint32_t boundary_cap;    /* 0: sized from the domain; 512000 = the reference cap (main.cpp:151) */
    const uint32_t* flags;    /* per particle (bit0 emitter, bit1 droplet; record slot 5.z,
                                 shaders/_struct_layouts.txt:35), or NULL for 0 */
} dfsph_scene;

/* Solver policy.  Replaces the fixed iteration loop of main.cpp:2766-2853.
 *   REFERENCE: zero-filled pressure, max(min_iter,1) Jacobi iterations with
 *              the reference's coefficient alpha/rho^2 and its boundary terms
 *              exactly as written; reproduces the reference step.
 *   CONVERGED: same operator and update, Jacobi coefficient rho^3*alpha/dt^k,
 *              boundary mass rho0*V_b in the alpha and pressure-force terms,
 *              iterates until the FP64 average compression error <= tolerance
 *              and no particle's residual exceeds residual_bound. */
typedef enum dfsph_mode { DFSPH_POLICY_REFERENCE = 0, DFSPH_POLICY_CONVERGED = 1 } dfsph_mode;
typedef enum dfsph_warm {
    DFSPH_WARM_ZERO = 0,      /* the reference's zero fill (main.cpp:2767-2769, :2823-2825) */
    DFSPH_WARM_PREVIOUS = 1,  /* previous step's converged pressure, scaled by min(dt_prev/dt, 1)^k */
    DFSPH_WARM_EXTERNAL = 2,  /* guesses set with dfsph_set_warm_start */
    DFSPH_WARM_MLP = 3,       /* learned per-particle guess (dfsph_mlp_*) for both solves */
    DFSPH_WARM_MLP_DENSITY = 4 /* learned guess for the density solve, the previous step's
                                  pressure (as DFSPH_WARM_PREVIOUS) for the divergence solve */
} dfsph_warm;

typedef struct dfsph_solver_policy {
    int32_t mode;             /* dfsph_mode */
    int32_t warm_start;       /* dfsph_warm */
    int32_t div_enabled;      /* main.cpp:2766 */
    int32_t min_iter_div;     /* 1 (main.cpp:261) */
    int32_t max_iter_div;     /* 100 (main.cpp:119) */
    int32_t min_iter_density; /* 2 (main.cpp:115) */
    int32_t max_iter_density; /* 100 (main.cpp:116) */
    float tol_div;            /* relative: 0.001 = 0.1 % */
    float tol_density;        /* relative: 0.001 = 0.1 % */
    float cfl;                /* dt = min(cfl * 2r / max|v|, dt_max) when dt == 0 */
    float dt_max;
    int32_t boundary_psi;     /* -1: by mode; 0: V_b as written; 1: rho0*V_b */
    int32_t strict;           /* 1: return DFSPH_E_NOT_CONVERGED when a solve hits max_iter */
    /* Every solve of the CONVERGED policy, whatever its start, converges only
     * once, besides the average criterion, no particle has a residual above
     * this bound (same units as the tolerance).  The average alone lets a few
     * over-pressured particles through; at the 0.1 % tolerance a zero start
     * leaves at most about 25 x the tolerance (dam breaks at 1M and 8.4M
     * particles), a warm start can leave far more, and with it a different
     * velocity field.  Much tighter tolerances need a looser bound or none.  A
     * warm-started solve that misses either criterion at max_iter is redone
     * from zero pressure (dfsph_step) or reported as not converged
     * (dfsph_batch_step).  0: default (25 x the solve's tolerance); < 0: no
     * bound. */
    float residual_bound;
    /* The divergence solve's start, when it should differ from warm_start's:
     * DFSPH_DIV_START_* (0 = as warm_start says).  The reference policy always
     * starts both solves from zero. */
    int32_t div_start;
    int32_t reserved[1];
} dfsph_solver_policy;
#define DFSPH_DIV_START_DEFAULT  0
#define DFSPH_DIV_START_ZERO     1
#define DFSPH_DIV_START_PREVIOUS 2
#define DFSPH_DIV_START_MLP      3

/* Per-step statistics (the reference shows only ImGui counters, main.cpp:3073-3077). */
typedef struct dfsph_step_stats {
    double dt;
    int32_t iters_div;
    int32_t iters_density;
    int32_t converged_div;
    int32_t converged_density;
    double err_div;           /* dt * mean complementarity residual of the last checked iterate (FP64) */
    double err_density;       /* mean over particles of |s - A p| where p > 0 and
                                 max(-(s - A p), 0) where p = 0, last checked iterate (FP64) */
    double max_speed;         /* max |v| at the start of the step */
    uint64_t nonfinite;       /* non-finite position/velocity components after the step */
    double gpu_ms;            /* device time of the step (HIP events on the context stream);
                                 without the next step's cell build, which dfsph_step enqueues
                                 after its last launch so it runs while the host returns */
    uint32_t launches;        /* kernel launches issued */
    uint32_t host_syncs;      /* host waits issued */
    int32_t warm_fallback;    /* 1: a warm start was unusable or left a residual above the
                                 bound, and the solve was redone from zero pressure */
    float max_residual_div;   /* largest per-particle residual of the last checked iterate */
    float max_residual_density;
    int32_t reserved;
} dfsph_step_stats;

/* Exportable fields, in original-index order unless stated. */
typedef enum dfsph_field {
    DFSPH_FIELD_POSITION = 0,           /* float3; particles_buf slot 0 (shaders/_struct_layouts.txt:21) */
    DFSPH_FIELD_VELOCITY = 1,           /* float3; slot 1 */
    DFSPH_FIELD_DENSITY = 2,            /* float; da.x (shaders/_struct_layouts.txt:51) */
    DFSPH_FIELD_ALPHA_OVER_RHO2 = 3,    /* float; da.y */
    DFSPH_FIELD_DENSITY_ADV = 4,        /* float; da.z */
    DFSPH_FIELD_DENSITY_CHANGE = 5,     /* float; da.w */
    DFSPH_FIELD_PRESSURE_DIV = 6,       /* float; final divergence-solve pressure */
    DFSPH_FIELD_PRESSURE = 7,           /* float; final density-solve pressure (pressure_a) */
    DFSPH_FIELD_PRESSURE_ACCEL = 8,     /* float3; last pressure acceleration */
    DFSPH_FIELD_ALPHA = 9,              /* float; alpha before division by rho^2 */
    DFSPH_FIELD_NEIGHBOR_COUNT = 10,    /* uint32; fluid j != i with q < 1 */
    DFSPH_FIELD_BOUNDARY_NEIGHBOR_COUNT = 11, /* uint32; boundary samples with q < 1 */
    DFSPH_FIELD_MORTON_KEY = 12,        /* uint32; reference Morton key of the step's positions */
    DFSPH_FIELD_SORTED_INDEX = 13,      /* uint32[N]; stable (key, index) order, sorted slot -> index */
    DFSPH_FIELD_CELL_STARTS = 14,       /* uint32[maxAxis^3 + 1]; exclusive scan with sentinel */
    DFSPH_FIELD_BOUNDARY_POSITION = 15, /* float3[Nb]; sample order */
    DFSPH_FIELD_BOUNDARY_VOLUME = 16,   /* float[Nb]; sample order */
    DFSPH_FIELD_ID = 21                 /* uint32; original index, in the library's sorted order */
} dfsph_field;

/* Derived constants (main.cpp:1994-2019, :2049-2105). */
typedef struct dfsph_derived {
    float h, cell, mass, knorm, gnorm, relax;
    uint32_t axes[3];         /* reference Morton axes (pow2) */
    uint32_t max_axis;
    uint32_t total_cells;     /* maxAxis^3 */
    float fine_cell;          /* FINE grid cell edge (h * max(fine_cell_scale, 1 + 2^-10)) */
    uint32_t fine_axes[3];
    /* IEEE-float equivalents of the reference's per-pair tests, so the device
     * decides neighbourhood from r^2 alone, bit-identically:                  */
    float r2_q1;              /* r^2 <  r2_q1     <=>  sqrt(r^2)/h < 1        */
    float r2_eps_gt;          /* r^2 >= r2_eps_gt <=>  sqrt(r^2) > 1e-7       */
    float r2_eps_lt;          /* r^2 <  r2_eps_lt <=>  sqrt(r^2) < 1e-7       */
} dfsph_derived;

/* ---- lifecycle: replaces createTierResources (main.cpp:1159-1282) ---- */
int dfsph_create(const dfsph_params* params, dfsph_ctx** out);
int dfsph_destroy(dfsph_ctx* ctx);
int dfsph_init_scene(dfsph_ctx* ctx, const dfsph_scene* scene);

/* ---- step: replaces the substep dispatch chain, main.cpp:2663-2867.
 * dt > 0 is used as given (the reference takes it from the wall clock,
 * main.cpp:2663); dt == 0 selects the CFL step of the policy. ---- */
int dfsph_step(dfsph_ctx* ctx, float dt, int32_t substeps, const dfsph_solver_policy* policy,
               dfsph_step_stats* stats);

/* ---- export: replaces the readbacks of capture_save (main.cpp:2438-2487) and
 * the Alembic pos/vel unpack (main.cpp:2875-2897). `bytes` must equal the
 * size reported by dfsph_field_size. ---- */
int dfsph_field_size(dfsph_ctx* ctx, int32_t field, size_t* bytes);
int dfsph_export(dfsph_ctx* ctx, int32_t field, void* dst, size_t bytes);

/* Initial pressure guesses for DFSPH_WARM_EXTERNAL, original-index order
 * (count = particle count); either pointer may be NULL for zero. */
int dfsph_set_warm_start(dfsph_ctx* ctx, const float* p_div, const float* p_density);

/* ---- captures: the sphWater F5/F9 capture of main.cpp:2438-2552, written as
 * StateWriter writes it (common/common-cpp/src/state_writer.cpp:13-87):
 * <root>/capture_NNNN/state.json plus particles.bin (128-byte records,
 * shaders/_struct_layouts.txt:11-42), density_alpha.bin, pressure.bin,
 * morton_codes.bin, sorted_index.bin and cell_starts.bin, all in
 * original-index order.  The cell-start table holds the intended exclusive
 * scan over the context's table (maxAxis^3 + 1 entries), not the reference's
 * 2^18-entry aliased one. ---- */
typedef struct dfsph_capture_info {
    uint64_t iteration;        /* meta.sphWater.iteration */
    uint32_t frame;            /* state.json "frame" (set by load / read_info) */
    uint32_t particle_count;   /* meta.sphWater.particle_count (set by load / read_info) */
    int32_t tier_index;        /* meta.sphWater.tier_index */
    int32_t preset_index;      /* meta.sphWater.preset_index */
    int32_t substeps;          /* meta.sphWater.substeps */
    int32_t reserved0;
    const char* camera_json;   /* save: a JSON object stored as meta.sphWater.camera, or NULL */
    int32_t reserved[4];
} dfsph_capture_info;
/* capture_save: writes <root_dir>/capture_<frame, 4 digits>; its path goes to
 * dir_out (dir_bytes including the terminator) when dir_out is not NULL.
 * Before the first step the per-step buffers are written as zeros. */
int dfsph_capture_save(dfsph_ctx* ctx, const char* root_dir, uint32_t frame, const dfsph_capture_info* info,
                       char* dir_out, size_t dir_bytes);
/* capture_load: checks the capture as main.cpp:2493-2527 does (sphWater meta,
 * particle_count, buffer sizes; DFSPH_E_IO otherwise), re-initialises the
 * scene from the records' positions, velocities and flags (boundary and
 * boundary_cap as in dfsph_scene), keeps pressure.bin as the previous
 * density pressure (DFSPH_WARM_PREVIOUS) and fills *info when not NULL. */
int dfsph_capture_load(dfsph_ctx* ctx, const char* capture_dir, int32_t boundary, uint32_t boundary_cap,
                       dfsph_capture_info* info);
/* The same checks without a context (host only). */
int dfsph_capture_read_info(const char* capture_dir, dfsph_capture_info* info);
/* StateReader::findLatest (common/common-cpp/src/state_reader.cpp:12-28): the
 * last capture_* directory of root_dir in name order. */
int dfsph_capture_find_latest(const char* root_dir, char* dir_out, size_t dir_bytes);

/* ---- Alembic particle export: replaces abc::ParticleWriter
 * (common/common-cpp/include/gpusims/alembic_writer.hpp:36-49,
 * common/common-cpp/src/alembic_writer.cpp:38-116) and the per-frame readback
 * and unpack of main.cpp:2875-2897.  Writes an Ogawa archive with one OPoints
 * object "particles" (.selfBnds, P, .pointIds, .velocities) on a uniform time
 * sampling of 1/fps, as the reference's writer does through the Alembic
 * library; frames stream to the file, the archive is complete after close. ---- */
typedef struct dfsph_abc dfsph_abc;
/* ParticleWriter::create(path, fps).  DFSPH_E_IO if the file cannot be made. */
int dfsph_abc_open(const char* path, double fps, dfsph_abc** out);
/* writeFrame: count >= 1 particles; positions and velocities xyz
 * interleaved (velocities may be NULL), ids NULL for 0..count-1. */
int dfsph_abc_write(dfsph_abc* writer, const float* positions, const float* velocities, const uint64_t* ids,
                    uint32_t count);
/* One frame of the context's state: positions and velocities in original
 * index order, ids 0..N-1 (main.cpp:2875-2897). */
int dfsph_abc_write_ctx(dfsph_abc* writer, dfsph_ctx* ctx);
/* Writes the group tables and headers and marks the archive complete; frees
 * the writer whatever the result. */
int dfsph_abc_close(dfsph_abc* writer);

/* ---- screen-space fluid render passes: replaces passes B-E of the frame
 * (main.cpp:2900-3132) -- the sphere-sprite depth pass
 * (shaders/particle_sprite.vert/frag.glsl), the bilateral depth smoothing
 * (shaders/bilateral_smooth.comp.glsl), the additive thickness pass
 * (shaders/thickness.vert/frag.glsl) and the composite
 * (shaders/composite.frag.glsl) -- over the context's particles, as compute
 * kernels that emulate the point-sprite rasterisation.  As the reference does:
 * every pixel of the depth image holds the sprite of the highest particle
 * index covering it (no depth test: the last draw wins), smoothing iterations
 * all run along x (the direction is packed once, main.cpp:2973), and the
 * composite scales the thickness by absorption_thickness_pad.y = 0
 * (main.cpp:2186).  Each of these can be changed through the fields below. ---- */
typedef struct dfsph_render_params {
    float view[16];             /* column-major, camera.view() (glm::lookAt) */
    float proj[16];             /* column-major, camera.projection() (Vulkan y flip) */
    float camera_pos[3];
    uint32_t width, height;     /* pixels, 1..16384 */
    float particle_radius;      /* 0: the context's radius */
    float point_scale;          /* height / (2 tan(fov / 2)), main.cpp:2135 */
    float thickness_per_particle;   /* THICKNESS_PER_PARTICLE_DEFAULT 0.005; 0..1 */
    int32_t bilateral_iterations;   /* 0..64; 4 */
    float sigma_spatial;        /* pixels; 3 */
    float sigma_depth;          /* NDC; 0.05 */
    int32_t bilateral_alternate;    /* 0: every iteration along x (the reference); 1: x, y, x, ... */
    float water_tint[3];        /* (0.10, 0.30, 0.50) */
    float roughness;            /* 0 */
    float sky_zenith[3];        /* (0.40, 0.55, 0.85) */
    float sky_horizon[3];       /* (0.85, 0.90, 0.95) */
    float exposure;             /* 1 */
    float fresnel_f0;           /* 0.02 */
    float absorption;           /* absorption_thickness_pad.x: thickness_per_particle in the reference */
    float thickness_scale;      /* absorption_thickness_pad.y: 0 in the reference (thickness unused) */
    uint32_t reserved[8];       /* zero */
} dfsph_render_params;
/* The reference's defaults (main.cpp:135-156, 274-282, 2133-2191) for a
 * camera at eye looking at target (gpusims::Camera orbit mode: lookAt with +y
 * up, perspective fov_deg over width/height, near 0.05, far 100, y flipped). */
int dfsph_render_defaults(const float eye[3], const float target[3], float fov_deg, uint32_t width,
                          uint32_t height, dfsph_render_params* out);
/* Render the context's particles.  Outputs (host, row 0 at the top, each may
 * be NULL): rgba width*height*4 floats (linear colour times exposure, alpha
 * 1), depth and smoothed width*height floats (NDC z, 1 = no particle) and
 * thickness width*height floats. */
int dfsph_render(dfsph_ctx* ctx, const dfsph_render_params* params, float* rgba, float* depth, float* smoothed,
                 float* thickness);

/* ---- emitters: the reserve-tail inject of shaders/apply_emitter.comp.glsl:29-65
 * (Emitter, main.cpp:237-244; at most EMITTER_CAP = 8, main.cpp:2597-2606).
 * New particle g (0 <= g < n_new) comes from emitter g % count into slot
 * s = particle count + g, at e.pos + e.radius * unit_ball_sample(s*7919 +
 * (g % count)*31) (PCG hash, :29-42), with e.velocity, flags 1 and original
 * index s.  Unlike the reference, whose particle count never grows (the
 * injected tail stays inactive), the new particles join the scene: the count
 * grows by *emitted = min(n_new, capacity - count).  Per-step fields are
 * exportable again after the next step. ---- */
#define DFSPH_MAX_EMITTERS 8
typedef struct dfsph_emitter {
    float pos[3];
    float radius;
    float velocity[3];
    float rate;        /* emissionRate, particles/s: carried; the caller chooses n_new */
} dfsph_emitter;
int dfsph_emit(dfsph_ctx* ctx, const dfsph_emitter* emitters, uint32_t count, uint32_t n_new,
               uint32_t* emitted);

/* ---- per-kernel timing: replaces GpuProfiler scopes around every dispatch
 * (common/common-cpp/src/gpu_profiler.cpp:55-132, main.cpp:2687-2859).  When
 * enabled, each launch on the context stream is bracketed by a HIP event pair;
 * launches that a converged solve skipped are not counted. ---- */
typedef struct dfsph_kernel_time {
    char name[32];
    uint64_t launches;
    double total_ms;
    double min_ms;
    double max_ms;
} dfsph_kernel_time;
/* enable: 0 off; n >= 1 times every n-th step (an event pair per launch costs
 * host time, so sampling keeps a timed run close to an untimed one).  Also
 * resets the counters. */
int dfsph_profile_enable(dfsph_ctx* ctx, int32_t enable);
int dfsph_profile_read(dfsph_ctx* ctx, dfsph_kernel_time* out, uint32_t cap, uint32_t* count);
/* Region marker for external profilers, standing in for the named
 * GpuProfiler scopes of gpu_profiler.cpp:55-132: launches one empty kernel,
 * k_trace_mark, on the context stream.  Two marks delimit a region in a
 * rocprofv3 kernel trace or counter run (dispatches between them, in
 * dispatch order); tag is the kernel argument and does nothing. */
int dfsph_trace_mark(dfsph_ctx* ctx, uint32_t tag);

/* ---- learned warm start (no reference counterpart: the reference zero-fills
 * both solves, main.cpp:2767-2769 and :2823-2825).  A per-particle network
 * 16 -> 64 -> 64 -> 64 -> 2 (ReLU, 9,538 FP32 parameters, BF16 MFMA inference
 * and training with FP32 accumulation) predicts the converged stiffness of the
 * divergence and density solves; DFSPH_WARM_MLP starts each solve from it and
 * the solver still iterates to tolerance.  Parameters are one flat FP32 vector
 * [W1 (16x64, row = input), b1, W2 (64x64), b2, W3 (64x64), b3, W4 (64x2), b4].
 * Dataset rows: 16 features, 2 targets (kappa_v, kappa_rho, dimensionless). ---- */
#define DFSPH_MLP_PARAMS 9538u
#define DFSPH_MLP_FEATURES 16u
#define DFSPH_MLP_OUTPUTS 2u

typedef struct dfsph_train_config {
    uint32_t epochs;
    uint32_t batch;           /* rows per optimiser step, multiple of 64 */
    float lr;                 /* Adam step size */
    float beta1, beta2, eps;
    uint64_t seed;            /* row shuffling */
    float val_fraction;       /* held-out rows (fixed hash split) */
    /* Expectile tau of the density output's loss: its squared error is weighted
     * 2 tau where the prediction is low and 2 (1 - tau) where high.  The solver
     * pays more for a low guess (the residual it leaves is smooth and Jacobi
     * removes smooth error slowly) than for a high one; 0 = 0.5 = plain MSE. */
    float expectile_density;
    uint32_t reserved[2];
} dfsph_train_config;

typedef struct dfsph_train_stats {
    double train_loss;        /* mean (expectile-weighted) squared error over the last epoch */
    double val_loss;          /* the same on the held-out rows */
    uint32_t steps;
    uint32_t train_rows;
    uint32_t val_rows;
    uint32_t reserved;
    double ms;
} dfsph_train_stats;

int dfsph_mlp_init(dfsph_ctx* ctx, uint64_t seed);                              /* He-uniform, zero biases */
int dfsph_mlp_set_params(dfsph_ctx* ctx, const float* theta, uint32_t count);   /* count = DFSPH_MLP_PARAMS */
int dfsph_mlp_get_params(dfsph_ctx* ctx, float* theta, uint32_t count);
/* Record (features, converged kappa) rows during subsequent dfsph_step calls:
 * 2 rows (one per solve) for one particle in `stride` per step, up to capacity. */
int dfsph_dataset_start(dfsph_ctx* ctx, uint32_t capacity_rows, uint32_t stride);
int dfsph_dataset_stop(dfsph_ctx* ctx, uint32_t* rows);
int dfsph_dataset_export(dfsph_ctx* ctx, float* features, float* targets, uint32_t rows);
int dfsph_dataset_import(dfsph_ctx* ctx, const float* features, const float* targets, uint32_t rows);
int dfsph_mlp_train(dfsph_ctx* ctx, const dfsph_train_config* cfg, dfsph_train_stats* stats);
/* Network outputs for caller rows (n x 16 -> n x 2), on the device. */
int dfsph_mlp_eval(dfsph_ctx* ctx, const float* features, uint32_t n, float* out);
/* Loss and its gradient (DFSPH_MLP_PARAMS values) for caller rows, n <= 65536. */
int dfsph_mlp_grad(dfsph_ctx* ctx, const float* features, const float* targets, uint32_t n,
                   float* grad, double* loss);
/* The same loss and gradient without allocating.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving or training loops, harnesses that forbid
 * hipMalloc).  The staged rows, the per-workgroup gradient partials and the
 * summed gradient live in one static device array of 2.79 MB per GPU, held
 * until the process ends, so n is limited to DFSPH_MLP_GRAD_NOALLOC_MAX_ROWS
 * (DFSPH_E_INVALID above it).  Results equal dfsph_mlp_grad's bit for bit.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (two contexts driven from two threads) race; calls from one thread do
 * not, since each returns after its stream has drained. */
#define DFSPH_MLP_GRAD_NOALLOC_MAX_ROWS 4096u
int dfsph_mlp_grad_noalloc(dfsph_ctx* ctx, const float* features, const float* targets, uint32_t n,
                           float* grad, double* loss);

/* ---- batched scenes.  No reference counterpart: the reference advances one
 * scene per process (frame loop main.cpp:2563-3158).  A batch packs many
 * independent scenes (each with its own dfsph_params: radius, tank, fluid)
 * into shared device arrays and advances all of them with ONE persistent
 * launch per substep -- the hash, sweeps and both Jacobi loops run on the
 * device until every scene meets its tolerance -- replayed from a captured
 * HIP graph.  Each scene's result is bit-identical to the same scene stepped
 * alone by dfsph_step with block_size 256 and the FINE grid. ---- */
typedef struct dfsph_batch dfsph_batch;

typedef struct dfsph_batch_config {
    uint32_t abi_version;     /* must be DFSPH_ABI_VERSION */
    int32_t device;           /* HIP device ordinal */
    uint32_t workgroups;      /* persistent grid (multiple of 8); 0 = every co-resident workgroup */
    uint32_t flags;           /* DFSPH_BATCH_* */
    uint32_t spin_limit;      /* polls a grid barrier waits before the substep gives up
                                 (DFSPH_E_DEVICE); 0 = default (about 0.3 s) */
    uint32_t reserved[3];     /* zero */
} dfsph_batch_config;

#define DFSPH_BATCH_NO_GRAPH   1u  /* launch directly each substep instead of replaying a graph */
#define DFSPH_BATCH_NO_XCD_MAP 2u  /* plain workgroup -> tile order (no per-XCD slabs) */

typedef struct dfsph_batch_info_t {
    uint32_t scenes, particles, boundary_particles, tiles, cells;
    uint32_t workgroups;      /* persistent grid in use */
    uint32_t max_workgroups;  /* co-resident limit of the device */
    uint32_t reserved;
} dfsph_batch_info_t;

/* params[s] must use the FINE grid; params[s].block_size and .flags are not
 * used.  scenes[s].count >= 1.  The boundary of each scene is sampled on its
 * own domain AABB. */
int dfsph_batch_create(const dfsph_batch_config* cfg, const dfsph_params* params,
                       const dfsph_scene* scenes, uint32_t count, dfsph_batch** out);
int dfsph_batch_destroy(dfsph_batch* batch);
int dfsph_batch_info(dfsph_batch* batch, dfsph_batch_info_t* out);
/* dfsph_trace_mark on the batch's stream. */
int dfsph_batch_trace_mark(dfsph_batch* batch, uint32_t tag);
/* One substep of every scene.  dt > 0: fixed for all scenes; 0: each scene's
 * own CFL step.  Warm start ZERO or PREVIOUS.  per_scene (count entries) and
 * total (iterations and errors: max over scenes; nonfinite: sum) may be NULL.
 * DFSPH_E_DEVICE if the persistent grid could not run co-resident (a grid
 * barrier waited spin_limit polls): that substep stopped between stages, and
 * the batch then refuses every further substep (destroy and re-create it). */
int dfsph_batch_step(dfsph_batch* batch, float dt, const dfsph_solver_policy* policy,
                     dfsph_step_stats* per_scene, dfsph_step_stats* total);
/* Fields of one scene, as dfsph_export (all but the reference-hash fields
 * MORTON_KEY, SORTED_INDEX and CELL_STARTS). */
int dfsph_batch_field_size(dfsph_batch* batch, uint32_t scene, int32_t field, size_t* bytes);
int dfsph_batch_export(dfsph_batch* batch, uint32_t scene, int32_t field, void* dst, size_t bytes);

/* ---- autotuner.  No reference counterpart (the reference's workgroup sizes
 * are fixed in its shaders).  Times `steps` substeps (after one untimed
 * substep) of every candidate configuration on the caller's scene(s) from the
 * same start and returns the fastest.  Single scene (FINE grid): block_size
 * 64/128/256 x neighbour masks / lists x XCD remap on/off x cell edge 1, 1.25
 * or 1.5 h (REFERENCE grid: block_size x XCD remap).  Batch: persistent grid
 * of 1/2/4/8 workgroups per CU (up to the co-resident limit) x per-XCD slabs
 * on/off.  cache_path (may be NULL): a text file of earlier choices keyed by
 * the device and every input the choice depends on (each scene's particle
 * count, radius, domain and constants, the grid kind, the caller's other
 * flags, the whole policy); a hit is returned without timing and a miss is
 * appended.  Returned flags are the caller's flags with the tuned bits
 * (DFSPH_FLAG_NEIGHBOR_LIST, DFSPH_FLAG_NO_XCD_REMAP; DFSPH_BATCH_NO_XCD_MAP)
 * replaced. ---- */
typedef struct dfsph_tune_result {
    uint32_t block_size;      /* single scene: dfsph_params.block_size */
    uint32_t flags;           /* single scene: dfsph_params.flags (DFSPH_FLAG_*) */
    uint32_t workgroups;      /* batch: dfsph_batch_config.workgroups */
    uint32_t batch_flags;     /* batch: dfsph_batch_config.flags */
    double ms_per_step;       /* mean substep time of the choice */
    double worst_ms_per_step; /* slowest candidate */
    uint32_t candidates;      /* configurations timed (0 on a cache hit) */
    int32_t cached;           /* 1: answered from cache_path */
    float fine_cell_scale;    /* single scene: dfsph_params.fine_cell_scale */
    uint32_t reserved;
} dfsph_tune_result;

int dfsph_autotune(const dfsph_params* params, const dfsph_scene* scene,
                   const dfsph_solver_policy* policy, uint32_t steps, const char* cache_path,
                   dfsph_tune_result* out);
int dfsph_batch_autotune(const dfsph_batch_config* cfg, const dfsph_params* params,
                         const dfsph_scene* scenes, uint32_t count, const dfsph_solver_policy* policy,
                         uint32_t steps, const char* cache_path, dfsph_tune_result* out);

/* Thread-local message of the last failing call on this thread. */
const char* dfsph_last_error(void);

/* ---- host utilities (no device needed) ---- */
int dfsph_default_policy(int32_t mode, dfsph_solver_policy* out);
int dfsph_derive(const dfsph_params* params, dfsph_derived* out);
/* Akinci boundary samples on the AABB (main.cpp:1332-1431): returns the count
 * through *count; writes min(count, cap_out) xyz triples when out != NULL. */
int dfsph_boundary_samples(const float dmin[3], const float dmax[3], float radius, uint32_t cap,
                           float* out, uint32_t cap_out, uint32_t* count);
/* Scene presets of main.cpp:191-232 (SphPreset, main.cpp:162-181):
 * 0 Dam-Break, 1 Central-Fountain, 2 Droplet-Impact, 3 Pour-from-Source. */
#define DFSPH_PRESET_COUNT 4
typedef struct dfsph_preset {
    char name[32];
    float brick_min[3], brick_max[3];       /* initial_brick_min / _max */
    int32_t add_droplet;
    float droplet_center[3];
    float droplet_radius;
    float droplet_velocity[3];              /* droplet_velocity_init */
    int32_t source_shape;                   /* EmitterShape: 0 none, 1 cylinder, 2 rectangle */
    float source_pos[3], source_size[3], source_velocity[3];
    float source_rate;                      /* source_emission_rate, particles/s */
    float domain_min[3], domain_max[3];
    float viscosity, cohesion, vorticity, roughness;
} dfsph_preset;
int dfsph_preset_get(int32_t index, dfsph_preset* out);
/* Initial fill of a preset (seedInitialFill, main.cpp:2266-2294, and
 * shaders/initial_fill.comp.glsl:22-65): a brick of
 * min(capacity, prod max(1, extent/2r)) particles at bmin + (ijk + 0.5) 2r,
 * then the droplet's min((4 dr^3)/3, capacity - brick) particles (dr =
 * max(1, droplet_radius/2r)) taken in x-fastest order from a (2 droplet_radius
 * / 2r)^3 lattice centred on the droplet, points beyond droplet_radius pulled
 * to 0.95 of it, with the droplet velocity and flags 2.  Counts go to *count
 * and *brick_count (either may be NULL); positions (count x 3), velocities
 * (count x 3) and flags (count) are written when not NULL. */
int dfsph_preset_fill(const dfsph_preset* preset, float radius, uint32_t capacity, float* positions,
                      float* velocities, uint32_t* flags, uint32_t* count, uint32_t* brick_count);
/* Seeded jittered lattice block: nx*ny*nz particles at spacing 2r starting at
 * origin + (i+0.5)*2r (shaders/initial_fill.comp.glsl:37-41), plus uniform
 * jitter of +-jitter*r per axis from a SplitMix64 counter hash of
 * (seed, id0 + id, axis).  Writes nx*ny*nz xyz triples. */
int dfsph_scene_block(uint32_t nx, uint32_t ny, uint32_t nz, float radius, const float origin[3],
                      uint64_t seed, float jitter, uint32_t id0, float* out);

#ifdef __cplusplus
}
#endif

#endif /* DFSPH_H */
