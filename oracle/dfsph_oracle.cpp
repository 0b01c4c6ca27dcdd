// dfsph_oracle.cpp -- CPU oracle for the DFSPH water substep.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product: only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the
// library built from it.  The shipped HIP path (gpu-sims_amd/) never links it.
//
// It restates, function by function, the reference's sph-water substep
// (StevenFAU/GPU-Sims, particle-fluids/sph-water):
//   * parameter derivation            src/main.cpp:1994-2019, :2021-2105
//   * Akinci boundary sampling         src/main.cpp:1332-1431
//   * CPU Morton counting sort         src/main.cpp:1436-1493
//   * boundary volume                  shaders/compute_boundary_volume.comp.glsl:55-92
//   * Morton key                       shaders/morton_code.comp.glsl:24-50
//   * cell count / scan / scatter      shaders/cell_count, prefix_sum_*, scatter (intent:
//                                      the stable counting sort of main.cpp:1452-1493)
//   * density and alpha                shaders/density_alpha.comp.glsl:99-179
//   * divergence source                shaders/compute_density_change.comp.glsl:82-151
//   * density source                   shaders/compute_density_adv.comp.glsl:80-153
//   * pressure acceleration            shaders/compute_pressure_accel.comp.glsl:84-157
//   * aij*p stencil                    shaders/compute_aij_pj.comp.glsl:89-163
//   * Jacobi updates                   shaders/jacobi_update_{density,divergence}.comp.glsl:47-59
//   * velocity apply                   shaders/apply_velocity.comp.glsl:45-51
//   * forces / position integration    shaders/integrate_forces.comp.glsl:65-143
//   * substep order                    src/main.cpp:2663-2867
//
// Each function is pinned by a check of its own in tests/test_oracle.py and
// tests/test_oracle_pins.py (golden vectors, closed forms on an exact lattice,
// symmetry, brute force on tiny inputs); only the emitter's hash constants are
// parity unpinned (the reference holds no emitter output).
//
// Everything is templated on the arithmetic type so the same text gives the
// FP32 oracle (bit-faithful evaluation order, built with -ffp-contract=off) and
// the FP64 oracle used to state per-field rounding error.
//
// Semantics fixed here where the reference is undefined (each is a documented
// deviation, not a guess at what a driver happens to do):
//   * sorted order: stable by (cell key, original index), exactly the order
//     main.cpp:1452-1493 produces; the GPU scatter's atomic order is not
//     reproducible (shaders/scatter.comp.glsl:30-32).
//   * cell_starts: the exclusive scan of the cell counts with the sentinel
//     cell_starts[cells] = N (main.cpp:1463-1485); the reference's GPU scan is
//     mis-wired (main.cpp:1826-1838).
//   * integrate_forces mode 0 reads every neighbour's velocity before any
//     particle's velocity is overwritten; the shader updates in place while
//     other invocations read (shaders/integrate_forces.comp.glsl:114,142).
//   * the Morton table has maxAxis^3 cells with no 2^18 cap (main.cpp:1294
//     caps it; the caps do not bind for 32^3 tables).
//
// New (no reference counterpart): the converged solver policy (Jacobi
// coefficient rho^3*alpha/dt^k, tolerance loop, FP64 error), warm starts, the
// CFL time step.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include <omp.h>

namespace {

// ---------------------------------------------------------------------------
// Public (C ABI) structs.  Kept layout-identical to oracle/oracle.py.
// ---------------------------------------------------------------------------
}  // namespace

extern "C" {

struct oracle_config {
    float r;            // particle radius
    float rho0;         // rest density
    float gravity[3];
    float viscosity;
    float cohesion;
    float vorticity;
    float dmin[3];
    float dmax[3];
    int32_t boundary;   // 1: sample the domain AABB with Akinci boundary particles
    uint32_t boundary_cap;  // 0: no cap; reference value 500*1024 (main.cpp:151)
};

struct oracle_derived {
    float h;            // support radius 4r          (main.cpp:269)
    float cell;         // Morton cell size 8r        (main.cpp:270)
    float mass;         // rho0 * (2r)^3              (main.cpp:2008-2011)
    float knorm;        // 8 / (pi h^3)               (main.cpp:2012-2015)
    float gnorm;        // 8 / (pi h^4)               (main.cpp:2016-2019)
    float relax;        // 0.5                        (main.cpp:122)
    uint32_t axes[3];   // pow2 cells per axis        (main.cpp:1994-2005)
    uint32_t max_axis;
    uint32_t total_cells;  // maxAxis^3               (main.cpp:2024)
};

struct oracle_policy {
    int32_t mode;             // 0 reference (fixed iterations, alpha/rho^2), 1 converged
    int32_t warm;             // 0 zero, 1 previous step's pressure, 2 caller-supplied
    int32_t div_enabled;      // main.cpp:2766
    int32_t min_iter_div;
    int32_t max_iter_div;
    int32_t min_iter_density;
    int32_t max_iter_density;
    float tol_div;            // relative, e.g. 0.001 = 0.1 %
    float tol_density;
    float cfl;                // dt = min(cfl * 2r / max|v|, dt_max) when dt == 0
    float dt_max;
    int32_t boundary_psi;     // boundary mass in the alpha and pressure-force terms:
                              // 0 = V_b as the reference writes it, 1 = rho0 * V_b
    float residual_bound;     // converged policy: largest per-particle residual also
                              // <= this (0: 25 x the tolerance, < 0: no bound)
};

struct oracle_stats {
    double dt;
    int32_t iters_div;
    int32_t iters_density;
    int32_t converged_div;
    int32_t converged_density;
    double err_div;           // dt * mean complementarity residual, last iterate checked
    double err_density;       // mean complementarity residual: |s - Ap| where p > 0,
                              // max(-(s - Ap), 0) where p = 0; last iterate checked
    double max_speed;
    uint64_t nonfinite;
};

}  // extern "C"

namespace {

constexpr float kAlphaEps = 1.0e-5f;   // main.cpp:121 / density_alpha.comp.glsl:74
constexpr float kJacobiRelax = 0.5f;   // main.cpp:122

// ---------------------------------------------------------------------------
// a1: parameter derivation
// ---------------------------------------------------------------------------
uint32_t next_pow2(uint32_t v) {       // main.cpp:1988-1993
    if (v <= 1) return 1u;
    v -= 1;
    v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16;
    return v + 1u;
}

oracle_derived derive(const oracle_config& c) {
    oracle_derived d{};
    d.h = c.r * 4.0f;                       // PARTICLE_RADIUS * SUPPORT_RADIUS_RATIO
    d.cell = c.r * 4.0f * 2.0f;             // ... * CELL_SIZE_RATIO_TO_SUPPORT
    const float spacing = 2.0f * c.r;
    d.mass = c.rho0 * spacing * spacing * spacing;
    const float pi = float(M_PI);
    d.knorm = 8.0f / (pi * d.h * d.h * d.h);
    d.gnorm = 8.0f / (pi * d.h * d.h * d.h * d.h);
    d.relax = kJacobiRelax;
    for (int a = 0; a < 3; ++a) {
        float ext = c.dmax[a] - c.dmin[a];
        uint32_t n = std::max(1u, uint32_t(std::ceil(ext / d.cell)));
        d.axes[a] = next_pow2(n);
    }
    d.max_axis = std::max({d.axes[0], d.axes[1], d.axes[2]});
    d.total_cells = d.max_axis * d.max_axis * d.max_axis;
    return d;
}

// ---------------------------------------------------------------------------
// a3: Morton key (morton_code.comp.glsl:24-36, main.cpp:1436-1445)
// ---------------------------------------------------------------------------
inline uint32_t expand_bits_10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
inline uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return (expand_bits_10(x) << 2) | (expand_bits_10(y) << 1) | expand_bits_10(z);
}

// GLSL clamp(x, 0, hi) then uint conversion.  NaN maps to cell 0 (fmax/fmin
// return the non-NaN operand), which is also what the HIP path does.
template <class R>
inline uint32_t cell_coord(R pos, R dmin, R cell, uint32_t axes) {
    R rel = (pos - dmin) / cell;
    R c = std::fmin(std::fmax(rel, R(0)), R(axes) - R(1));
    return uint32_t(c);
}

// ---------------------------------------------------------------------------
// a2: cubic spline (density_alpha.comp.glsl:87-97)
// ---------------------------------------------------------------------------
template <class R>
inline R kernel_W(R q, R k) {
    if (q < R(0.5)) return k * (R(6) * q * q * q - R(6) * q * q + R(1));
    else if (q < R(1)) { R omq = R(1) - q; return k * R(2) * omq * omq * omq; }
    return R(0);
}
template <class R>
inline R kernel_gradW_scale(R r_mag, R q, R g) {
    R poly = R(0);
    if (q < R(0.5)) poly = R(18) * q * q - R(12) * q;
    else if (q < R(1)) { R omq = R(1) - q; poly = R(-6) * omq * omq; }
    return g * poly / r_mag;
}

// ---------------------------------------------------------------------------
// a17: Akinci boundary sampling (main.cpp:1332-1431), float arithmetic as glm.
// ---------------------------------------------------------------------------
std::vector<float> boundary_samples(const float dmin[3], const float dmax[3], float r,
                                    uint32_t cap) {
    std::vector<float> out;
    const float spacing = 2.0f * r;
    if (spacing <= 0.0f) return out;
    auto push = [&](const float p[3]) -> bool {
        if (cap && out.size() / 3 >= cap) return false;
        out.push_back(p[0]); out.push_back(p[1]); out.push_back(p[2]);
        return true;
    };
    auto sample_face = [&](int na, float plane, int au, int av) {
        float u_lo = dmin[au] + spacing, u_hi = dmax[au] - spacing;
        float v_lo = dmin[av] + spacing, v_hi = dmax[av] - spacing;
        if (u_lo > u_hi || v_lo > v_hi) return;
        int n_v = int(std::floor((v_hi - v_lo) / spacing));
        for (int j = 0; j <= n_v; ++j) {
            float v = v_lo + float(j) * spacing;
            if (v > v_hi + 1e-6f) break;
            float offset = (j & 1) ? 0.5f * spacing : 0.0f;
            int n_u = int(std::floor((u_hi - u_lo - offset) / spacing));
            for (int i = 0; i <= n_u; ++i) {
                float u = u_lo + offset + float(i) * spacing;
                if (u > u_hi + 1e-6f) break;
                float p[3] = {0.0f, 0.0f, 0.0f};
                p[na] = plane; p[au] = u; p[av] = v;
                if (!push(p)) return;
            }
        }
    };
    sample_face(1, dmin[1], 0, 2);
    sample_face(1, dmax[1], 0, 2);
    sample_face(0, dmin[0], 1, 2);
    sample_face(0, dmax[0], 1, 2);
    sample_face(2, dmin[2], 0, 1);
    sample_face(2, dmax[2], 0, 1);
    auto sample_edge = [&](int fa, float p0, float p1, float p2) {
        float t_lo = dmin[fa] + spacing, t_hi = dmax[fa] - spacing;
        if (t_lo > t_hi) return;
        int n = int(std::floor((t_hi - t_lo) / spacing));
        for (int i = 0; i <= n; ++i) {
            float t = t_lo + float(i) * spacing;
            if (t > t_hi + 1e-6f) break;
            float p[3] = {p0, p1, p2};
            p[fa] = t;
            if (!push(p)) return;
        }
    };
    sample_edge(0, 0.0f, dmin[1], dmin[2]);
    sample_edge(0, 0.0f, dmin[1], dmax[2]);
    sample_edge(0, 0.0f, dmax[1], dmin[2]);
    sample_edge(0, 0.0f, dmax[1], dmax[2]);
    sample_edge(1, dmin[0], 0.0f, dmin[2]);
    sample_edge(1, dmin[0], 0.0f, dmax[2]);
    sample_edge(1, dmax[0], 0.0f, dmin[2]);
    sample_edge(1, dmax[0], 0.0f, dmax[2]);
    sample_edge(2, dmin[0], dmin[1], 0.0f);
    sample_edge(2, dmin[0], dmax[1], 0.0f);
    sample_edge(2, dmax[0], dmin[1], 0.0f);
    sample_edge(2, dmax[0], dmax[1], 0.0f);
    const float corners[8][3] = {
        {dmin[0], dmin[1], dmin[2]}, {dmax[0], dmin[1], dmin[2]},
        {dmin[0], dmax[1], dmin[2]}, {dmax[0], dmax[1], dmin[2]},
        {dmin[0], dmin[1], dmax[2]}, {dmax[0], dmin[1], dmax[2]},
        {dmin[0], dmax[1], dmax[2]}, {dmax[0], dmax[1], dmax[2]}};
    for (const auto& c : corners) push(c);
    return out;
}

// ---------------------------------------------------------------------------
// a3-a7: keys plus the stable counting sort of main.cpp:1452-1493.
// ---------------------------------------------------------------------------
template <class R>
void hash_grid(uint32_t n, const R* x, const R dmin[3], R cell, const uint32_t axes[3],
               uint32_t total_cells, std::vector<uint32_t>& keys,
               std::vector<uint32_t>& sorted, std::vector<uint32_t>& starts) {
    keys.assign(n, 0u);
    sorted.assign(n, 0u);
    starts.assign(size_t(total_cells) + 1, 0u);
    std::vector<uint32_t> counts(total_cells, 0u);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t cx = cell_coord(x[3 * i + 0], dmin[0], cell, axes[0]);
        uint32_t cy = cell_coord(x[3 * i + 1], dmin[1], cell, axes[1]);
        uint32_t cz = cell_coord(x[3 * i + 2], dmin[2], cell, axes[2]);
        uint32_t m = morton3(cx, cy, cz);
        keys[i] = m;
        counts[m]++;
    }
    uint32_t run = 0;
    for (uint32_t c = 0; c < total_cells; ++c) { starts[c] = run; run += counts[c]; }
    starts[total_cells] = run;
    std::vector<uint32_t> cursor(starts.begin(), starts.begin() + total_cells);
    for (uint32_t i = 0; i < n; ++i) sorted[cursor[keys[i]]++] = i;
}

// ---------------------------------------------------------------------------
// Simulation state (original index order, as in the reference's buffers).
// ---------------------------------------------------------------------------
struct SimBase {
    virtual ~SimBase() = default;
    virtual void prepare(int psi) = 0;
    virtual void apply_A(int mode_k, double dt, int psi, const double* p, double* ap, double* acc) = 0;
    virtual void source(int adv, double dt) = 0;
    virtual void forces(double dt) = 0;
    virtual int step(float dt_in, const oracle_policy& pol, oracle_stats* st) = 0;
    virtual int get(int field, double* out) const = 0;
    virtual int get_u32(int field, uint32_t* out) const = 0;
    virtual int set(int field, const double* in) = 0;
    virtual uint32_t count(int which) const = 0;
    virtual oracle_derived derived() const = 0;
};

enum Field {
    F_X = 0, F_V = 1, F_RHO = 2, F_AOR2 = 3, F_RHO_ADV = 4, F_RHO_CHANGE = 5,
    F_P_DIV = 6, F_P_DENSITY = 7, F_ACCEL = 8, F_ALPHA = 9, F_NBR = 10, F_NBR_B = 11,
    F_KEYS = 12, F_SORTED = 13, F_CELL_STARTS = 14, F_BX = 15, F_BV = 16,
    F_WARM_DIV = 17, F_WARM_DENSITY = 18, F_GF = 19, F_GB = 20,
    F_RES_DIV = 22, F_RES_DENSITY = 23,
};

template <class R>
struct Sim final : SimBase {
    oracle_config cfg{};
    oracle_derived d{};
    // arithmetic-type copies of the parameters
    R h, cell, mass, knorm, gnorm, rho0, relax, visc, coh, vort;
    R g[3], dmin[3], dmax[3];
    uint32_t n = 0, nb = 0;
    std::vector<R> x, v, v_tmp, da, pa, pb, accel, aij, alpha, gf, gb;
    std::vector<R> p_div_final, p_den_final, warm_div, warm_den;
    std::vector<uint32_t> nbr, nbr_b;
    std::vector<uint32_t> keys, sorted, starts;
    std::vector<R> bx, bvol;
    std::vector<uint32_t> bkeys, bsorted, bstarts;
    // Per-particle complementarity residuals of the last iterate each solve
    // checked (the terms whose mean is the solve's FP64 error).
    std::vector<double> res_div, res_den;
    double dt_prev = 0.0;
    bool have_prev = false;

    Sim(const oracle_config& c, uint32_t n_, const float* x_in, const float* v_in) : cfg(c) {
        d = derive(c);
        h = R(d.h); cell = R(d.cell); mass = R(d.mass); knorm = R(d.knorm); gnorm = R(d.gnorm);
        rho0 = R(c.rho0); relax = R(d.relax); visc = R(c.viscosity); coh = R(c.cohesion);
        vort = R(c.vorticity);
        for (int a = 0; a < 3; ++a) { g[a] = R(c.gravity[a]); dmin[a] = R(c.dmin[a]); dmax[a] = R(c.dmax[a]); }
        n = n_;
        x.resize(3 * size_t(n)); v.assign(3 * size_t(n), R(0));
        for (size_t i = 0; i < 3 * size_t(n); ++i) {
            x[i] = R(x_in[i]);
            if (v_in) v[i] = R(v_in[i]);
        }
        v_tmp.resize(3 * size_t(n));
        da.assign(4 * size_t(n), R(0));
        pa.assign(n, R(0)); pb.assign(n, R(0));
        accel.assign(3 * size_t(n), R(0)); aij.assign(n, R(0));
        alpha.assign(n, R(0));
        gf.assign(3 * size_t(n), R(0)); gb.assign(3 * size_t(n), R(0));
        p_div_final.assign(n, R(0)); p_den_final.assign(n, R(0));
        warm_div.assign(n, R(0)); warm_den.assign(n, R(0));
        nbr.assign(n, 0u); nbr_b.assign(n, 0u);
        if (c.boundary) init_boundary();
    }

    oracle_derived derived() const override { return d; }

    // Hash and density/alpha of the current positions, without stepping.
    void prepare(int psi) override {
        hash_grid<R>(n, x.data(), dmin, cell, d.axes, d.total_cells, keys, sorted, starts);
        density_alpha(psi ? rho0 : R(1));
    }
    // (A p)_i of the pressure operator: pressure_accel(p) then aij_pj(mode).
    void apply_A(int mode_k, double dt, int psi, const double* p, double* ap, double* acc) override {
        std::vector<R> pv(n);
        for (uint32_t i = 0; i < n; ++i) pv[i] = R(p[i]);
        pressure_accel(pv, psi ? rho0 : R(1));
        aij_pj(mode_k ? 1 : 0, R(dt));
        for (uint32_t i = 0; i < n; ++i) ap[i] = double(aij[i]);
        if (acc) for (size_t i = 0; i < accel.size(); ++i) acc[i] = double(accel[i]);
    }
    // One source pass of the current state after prepare(): divergence
    // source (adv = 0 -> da.w) or density source (adv = 1 -> da.z).
    void source(int adv, double dt) override { density_source(adv ? 1 : 0, R(dt)); }
    // One force pass of the current state after prepare(): v <- v + dt (g +
    // viscosity + cohesion).
    void forces(double dt) override { integrate_forces(R(dt)); }
    uint32_t count(int which) const override { return which == 0 ? n : nb; }

    // main.cpp:2306-2372 plus compute_boundary_volume.comp.glsl:55-92
    void init_boundary() {
        std::vector<float> bp = boundary_samples(cfg.dmin, cfg.dmax, cfg.r, cfg.boundary_cap);
        nb = uint32_t(bp.size() / 3);
        bx.resize(bp.size());
        for (size_t i = 0; i < bp.size(); ++i) bx[i] = R(bp[i]);
        hash_grid<R>(nb, bx.data(), dmin, cell, d.axes, d.total_cells, bkeys, bsorted, bstarts);
        bvol.assign(nb, R(0));
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < int64_t(nb); ++ii) {
            uint32_t i = uint32_t(ii);
            R pi[3] = {bx[3 * i], bx[3 * i + 1], bx[3 * i + 2]};
            R delta = kernel_W(R(0), knorm);
            walk(pi, bstarts, bsorted, [&](uint32_t j) {
                if (j == i) return;
                R rij[3] = {pi[0] - bx[3 * j], pi[1] - bx[3 * j + 1], pi[2] - bx[3 * j + 2]};
                R r = std::sqrt(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
                R q = r / h;
                if (q >= R(1)) return;
                delta += kernel_W(q, knorm);
            });
            bvol[i] = R(1) / std::max(delta, R(1.0e-9f));
        }
    }

    // 27-cell walk in dz, dy, dx order (density_alpha.comp.glsl:112-122).
    template <class F>
    void walk(const R pi[3], const std::vector<uint32_t>& st, const std::vector<uint32_t>& so,
              F&& fn) const {
        int ci[3];
        for (int a = 0; a < 3; ++a) ci[a] = int(cell_coord(pi[a], dmin[a], cell, d.axes[a]));
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int nc[3] = {ci[0] + dx, ci[1] + dy, ci[2] + dz};
                    if (nc[0] < 0 || nc[1] < 0 || nc[2] < 0) continue;
                    if (nc[0] >= int(d.axes[0]) || nc[1] >= int(d.axes[1]) || nc[2] >= int(d.axes[2]))
                        continue;
                    uint32_t m = morton3(uint32_t(nc[0]), uint32_t(nc[1]), uint32_t(nc[2]));
                    for (uint32_t k = st[m]; k < st[m + 1]; ++k) fn(so[k]);
                }
    }

    inline void pos(uint32_t i, R p[3]) const { p[0] = x[3 * i]; p[1] = x[3 * i + 1]; p[2] = x[3 * i + 2]; }

    // density_alpha.comp.glsl:99-179 (also records the fluid/boundary gradient
    // sums and neighbour counts for the checks).
    void density_alpha(R bscale) {
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < int64_t(n); ++ii) {
            uint32_t i = uint32_t(ii);
            R pi[3]; pos(i, pi);
            R density = R(0), sg[3] = {0, 0, 0}, sgsq = R(0), sgf[3] = {0, 0, 0};
            uint32_t cnt = 0, cntb = 0;
            walk(pi, starts, sorted, [&](uint32_t j) {
                R rij[3] = {pi[0] - x[3 * j], pi[1] - x[3 * j + 1], pi[2] - x[3 * j + 2]};
                R r = std::sqrt(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
                R q = r / h;
                if (q >= R(1)) return;
                density += mass * kernel_W(q, knorm);
                if (j != i) cnt++;
                if (j != i && r > R(1e-7f)) {
                    R s = kernel_gradW_scale(r, q, gnorm);
                    R gt[3] = {mass * (s * rij[0]), mass * (s * rij[1]), mass * (s * rij[2])};
                    for (int a = 0; a < 3; ++a) { sg[a] += gt[a]; sgf[a] += gt[a]; }
                    sgsq += gt[0] * gt[0] + gt[1] * gt[1] + gt[2] * gt[2];
                }
            });
            R sgb[3] = {0, 0, 0};
            if (nb > 0) {
                walk(pi, bstarts, bsorted, [&](uint32_t j) {
                    R rij[3] = {pi[0] - bx[3 * j], pi[1] - bx[3 * j + 1], pi[2] - bx[3 * j + 2]};
                    R r = std::sqrt(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
                    R q = r / h;
                    if (q >= R(1)) return;
                    cntb++;
                    density += rho0 * bvol[j] * kernel_W(q, knorm);
                    if (r > R(1e-7f)) {
                        R s = kernel_gradW_scale(r, q, gnorm);
                        for (int a = 0; a < 3; ++a) {
                            R gw = s * rij[a];
                            sg[a] += (bscale * bvol[j]) * gw;
                            sgb[a] += bvol[j] * gw;
                        }
                    }
                });
            }
            R denom = std::max(sg[0] * sg[0] + sg[1] * sg[1] + sg[2] * sg[2] + sgsq, R(kAlphaEps));
            R al = R(1) / denom;
            R stored = al / std::max(density * density, R(kAlphaEps));
            da[4 * i + 0] = density; da[4 * i + 1] = stored; da[4 * i + 2] = R(0); da[4 * i + 3] = R(0);
            alpha[i] = al;
            nbr[i] = cnt; nbr_b[i] = cntb;
            for (int a = 0; a < 3; ++a) { gf[3 * i + a] = sgf[a]; gb[3 * i + a] = sgb[a]; }
        }
    }

    // compute_density_change.comp.glsl:82-151 (mode 1 -> da.w) and
    // compute_density_adv.comp.glsl:80-153 (mode 0 -> da.z).
    void density_source(int adv, R dt) {
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < int64_t(n); ++ii) {
            uint32_t i = uint32_t(ii);
            R pi[3]; pos(i, pi);
            R vi[3] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
            R rho_i = da[4 * i];
            R delta = R(0);
            walk(pi, starts, sorted, [&](uint32_t j) {
                if (j == i) return;
                R rij[3] = {pi[0] - x[3 * j], pi[1] - x[3 * j + 1], pi[2] - x[3 * j + 2]};
                R r = std::sqrt(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
                R q = r / h;
                if (q >= R(1) || r < R(1e-7f)) return;
                R s = kernel_gradW_scale(r, q, gnorm);
                R dv[3] = {vi[0] - v[3 * j], vi[1] - v[3 * j + 1], vi[2] - v[3 * j + 2]};
                delta += dv[0] * (s * rij[0]) + dv[1] * (s * rij[1]) + dv[2] * (s * rij[2]);
            });
            R V_i = mass / std::max(rho_i, R(kAlphaEps));
            // density_adv sums the boundary terms separately and adds dt*sum
            // (compute_density_adv.comp.glsl:124-150); density_change adds each
            // term straight into the result (compute_density_change.comp.glsl:124).
            R out = adv ? rho_i / std::max(rho0, R(1e-7f)) + dt * V_i * delta : V_i * delta;
            R db = R(0);
            if (nb > 0) {
                walk(pi, bstarts, bsorted, [&](uint32_t j) {
                    R rij[3] = {pi[0] - bx[3 * j], pi[1] - bx[3 * j + 1], pi[2] - bx[3 * j + 2]};
                    R r = std::sqrt(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
                    R q = r / h;
                    if (q >= R(1) || r < R(1e-7f)) return;
                    R s = kernel_gradW_scale(r, q, gnorm);
                    R t = bvol[j] * (vi[0] * (s * rij[0]) + vi[1] * (s * rij[1]) + vi[2] * (s * rij[2]));
                    if (adv) db += t; else out += t;
                });
                if (adv) out += dt * db;
            }
            da[4 * i + (adv ? 2 : 3)] = out;
        }
    }

    // compute_pressure_accel.comp.glsl:84-157
    void pressure_accel(const std::vector<R>& p, R bscale) {
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < int64_t(n); ++ii) {
            uint32_t i = uint32_t(ii);
            R pi[3]; pos(i, pi);
            R rho_i = da[4 * i];
            R pr_i = p[i] / std::max(rho_i * rho_i, R(kAlphaEps));
            R a[3] = {0, 0, 0};
            walk(pi, starts, sorted, [&](uint32_t j) {
                if (j == i) return;
                R rij[3] = {pi[0] - x[3 * j], pi[1] - x[3 * j + 1], pi[2] - x[3 * j + 2]};
                R r = std::sqrt(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
                R q = r / h;
                if (q >= R(1) || r < R(1e-7f)) return;
                R s = kernel_gradW_scale(r, q, gnorm);
                R rho_j = da[4 * j];
                R pr_j = p[j] / std::max(rho_j * rho_j, R(kAlphaEps));
                R f = mass * (pr_i + pr_j);
                for (int k = 0; k < 3; ++k) a[k] -= f * (s * rij[k]);
            });
            if (nb > 0 && std::fabs(pr_i) > R(kAlphaEps)) {
                walk(pi, bstarts, bsorted, [&](uint32_t j) {
                    R rij[3] = {pi[0] - bx[3 * j], pi[1] - bx[3 * j + 1], pi[2] - bx[3 * j + 2]};
                    R r = std::sqrt(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
                    R q = r / h;
                    if (q >= R(1) || r < R(1e-7f)) return;
                    R s = kernel_gradW_scale(r, q, gnorm);
                    for (int k = 0; k < 3; ++k) a[k] -= pr_i * (bscale * bvol[j]) * (s * rij[k]);
                });
            }
            for (int k = 0; k < 3; ++k) accel[3 * i + k] = a[k];
        }
    }

    // compute_aij_pj.comp.glsl:89-163; mode 0 (density) scales by dt^2, mode 1 by dt.
    void aij_pj(int mode, R dt) {
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < int64_t(n); ++ii) {
            uint32_t i = uint32_t(ii);
            R pi[3]; pos(i, pi);
            R rho_i = da[4 * i];
            R ai[3] = {accel[3 * i], accel[3 * i + 1], accel[3 * i + 2]};
            R sum = R(0);
            walk(pi, starts, sorted, [&](uint32_t j) {
                if (j == i) return;
                R rij[3] = {pi[0] - x[3 * j], pi[1] - x[3 * j + 1], pi[2] - x[3 * j + 2]};
                R r = std::sqrt(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
                R q = r / h;
                if (q >= R(1) || r < R(1e-7f)) return;
                R s = kernel_gradW_scale(r, q, gnorm);
                R da_[3] = {ai[0] - accel[3 * j], ai[1] - accel[3 * j + 1], ai[2] - accel[3 * j + 2]};
                sum += da_[0] * (s * rij[0]) + da_[1] * (s * rij[1]) + da_[2] * (s * rij[2]);
            });
            R V_i = mass / std::max(rho_i, R(kAlphaEps));
            sum *= V_i;
            if (nb > 0) {
                R sb = R(0);
                walk(pi, bstarts, bsorted, [&](uint32_t j) {
                    R rij[3] = {pi[0] - bx[3 * j], pi[1] - bx[3 * j + 1], pi[2] - bx[3 * j + 2]};
                    R r = std::sqrt(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
                    R q = r / h;
                    if (q >= R(1) || r < R(1e-7f)) return;
                    R s = kernel_gradW_scale(r, q, gnorm);
                    sb += bvol[j] * (ai[0] * (s * rij[0]) + ai[1] * (s * rij[1]) + ai[2] * (s * rij[2]));
                });
                sum += sb;
            }
            if (mode == 0) sum *= dt * dt; else sum *= dt;
            aij[i] = sum;
        }
    }

    // apply_velocity.comp.glsl:45-51
    void apply_velocity(R dt) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < int64_t(n); ++i)
            for (int k = 0; k < 3; ++k) v[3 * i + k] = v[3 * i + k] + dt * accel[3 * i + k];
    }

    // integrate_forces.comp.glsl:99-143 (mode 0), race-free reads.
    void integrate_forces(R dt) {
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < int64_t(n); ++ii) {
            uint32_t i = uint32_t(ii);
            R pi[3]; pos(i, pi);
            R vi[3] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
            R av[3] = {0, 0, 0}, ac[3] = {0, 0, 0};
            walk(pi, starts, sorted, [&](uint32_t j) {
                if (j == i) return;
                R rij[3] = {pi[0] - x[3 * j], pi[1] - x[3 * j + 1], pi[2] - x[3 * j + 2]};
                R r = std::sqrt(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
                R q = r / h;
                if (q >= R(1) || r < R(1e-7f)) return;
                R rho_j = std::max(da[4 * j], R(1e-3f));
                R W = kernel_W(q, knorm);
                R fv = visc * (mass / rho_j);
                R fc = coh * mass * W;
                for (int k = 0; k < 3; ++k) {
                    av[k] += fv * (v[3 * j + k] - vi[k]) * W;
                    ac[k] -= fc * (rij[k] / r);
                }
            });
            for (int k = 0; k < 3; ++k) {
                R at = g[k] + av[k] + ac[k] + R(0) * vort;
                v_tmp[3 * i + k] = vi[k] + dt * at;
            }
        }
        v.swap(v_tmp);
    }

    // integrate_forces.comp.glsl:65-100 (mode 1)
    void integrate_position(R dt) {
        R margin = R(0.125f) * h;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < int64_t(n); ++i) {
            for (int k = 0; k < 3; ++k) {
                R pn = x[3 * i + k] + dt * v[3 * i + k];
                R vn = v[3 * i + k];
                if (pn < dmin[k] - margin) { pn = dmin[k]; vn = std::max(vn, R(0)); }
                else if (pn > dmax[k] + margin) { pn = dmax[k]; vn = std::min(vn, R(0)); }
                x[3 * i + k] = pn; v[3 * i + k] = vn;
            }
        }
    }

    // One pressure solve.  mode_k: 1 divergence (scale dt, source -rho_dot),
    // 0 density (scale dt^2, source 1 - rho_adv).
    void solve(int mode_k, R dt, const oracle_policy& pol, const std::vector<R>& init,
               int& iters, int& conv, double& err_out, std::vector<R>& p_final) {
        const bool reference = pol.mode == 0;
        const int min_it = mode_k ? pol.min_iter_div : pol.min_iter_density;
        const int max_it = mode_k ? pol.max_iter_div : pol.max_iter_density;
        const double tol = mode_k ? pol.tol_div : pol.tol_density;
        const double bound = pol.residual_bound == 0.0f ? 25.0f * float(tol) : pol.residual_bound;
        const R bscale = pol.boundary_psi ? rho0 : R(1);
        // Zero-fill (main.cpp:2767-2772 / :2823-2828) or warm start.
        std::fill(accel.begin(), accel.end(), R(0));
        for (uint32_t i = 0; i < n; ++i) { pa[i] = reference ? R(0) : init[i]; pb[i] = R(0); }
        std::vector<R> coef(n);
        for (uint32_t i = 0; i < n; ++i) {
            if (reference) coef[i] = da[4 * i + 1];
            else {
                R rho = da[4 * i];
                R c = alpha[i] * rho * rho * rho;
                coef[i] = mode_k ? c / dt : c / (dt * dt);
            }
        }
        std::vector<double> e(n);
        std::vector<R>* rd = &pa;
        std::vector<R>* wr = &pb;
        int it = 0;
        double err = 0.0;
        conv = 0;
        const int fixed = std::max(reference ? min_it : 1, 1);
        for (;;) {
            pressure_accel(*rd, bscale);
            aij_pj(mode_k ? 1 : 0, dt);
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < int64_t(n); ++i) {
                R s = mode_k ? -da[4 * i + 3] : R(1) - da[4 * i + 2];
                R res = s - aij[i];
                // complementarity residual: |r| where p > 0, compression only where p = 0
                e[i] = (*rd)[i] > R(0) ? std::fabs(double(res)) : std::max(-double(res), 0.0);
                (*wr)[i] = std::max((*rd)[i] - relax * (s - aij[i]) * coef[i], R(0));
            }
            long double acc = 0.0L;
            double emax = 0.0;
            for (uint32_t i = 0; i < n; ++i) { acc += e[i]; emax = std::max(emax, e[i]); }
            err = n ? double(acc / (long double)n) : 0.0;
            if (mode_k) { err *= double(dt); emax *= double(dt); }
            std::swap(rd, wr);
            ++it;
            if (reference) { if (it >= fixed) break; }
            else {
                const bool bounded = !(bound > 0.0) || emax <= bound;
                if (err <= tol && it >= std::max(min_it, 1) && bounded) { conv = 1; break; }
                if (it >= max_it) break;
            }
        }
        if (reference) conv = 1;
        iters = it;
        err_out = err;
        (mode_k ? res_div : res_den) = e;
        // Final accel from the last written pressure, then v += dt a
        // (main.cpp:2793-2797 / :2849-2853).
        pressure_accel(*rd, bscale);
        apply_velocity(dt);
        p_final = *rd;
    }

    int step(float dt_in, const oracle_policy& pol, oracle_stats* st) override {
        // Time step: explicit, or CFL on the speed at the start of the step.
        float vmax2 = 0.0f;
        for (uint32_t i = 0; i < n; ++i) {
            float vx = float(v[3 * i]), vy = float(v[3 * i + 1]), vz = float(v[3 * i + 2]);
            float s2 = vx * vx + vy * vy + vz * vz;
            if (s2 > vmax2) vmax2 = s2;
        }
        float vmax = std::sqrt(vmax2);
        float dtf = dt_in;
        if (!(dtf > 0.0f)) {
            dtf = pol.dt_max;
            if (vmax > 0.0f) dtf = std::min(pol.cfl * (2.0f * cfg.r) / vmax, pol.dt_max);
        }
        const R dt = R(dtf);

        // Hash (main.cpp:2693-2744 with the intended semantics).
        hash_grid<R>(n, x.data(), dmin, cell, d.axes, d.total_cells, keys, sorted, starts);
        density_alpha(pol.boundary_psi ? rho0 : R(1));
        density_source(0, dt);   // da.w

        // Warm-start guesses (previous step's pressure scaled by the dt ratio,
        // never scaled up).
        std::vector<R> init_div(n, R(0)), init_den(n, R(0));
        if (pol.warm == 1 && have_prev) {
            R ratio = R(std::min(dt_prev / double(dt), 1.0));
            for (uint32_t i = 0; i < n; ++i) {
                init_div[i] = p_div_final[i] * ratio;
                init_den[i] = p_den_final[i] * ratio * ratio;
            }
        } else if (pol.warm == 2) {
            init_div = warm_div;
            init_den = warm_den;
        }

        int iv = 0, cv = 1, id = 0, cd = 1;
        double ev = 0.0, ed = 0.0;
        if (pol.div_enabled) solve(1, dt, pol, init_div, iv, cv, ev, p_div_final);
        integrate_forces(dt);
        density_source(1, dt);   // da.z
        solve(0, dt, pol, init_den, id, cd, ed, p_den_final);
        integrate_position(dt);
        dt_prev = double(dt);
        have_prev = true;

        uint64_t bad = 0;
        for (size_t i = 0; i < x.size(); ++i)
            if (!std::isfinite(double(x[i])) || !std::isfinite(double(v[i]))) ++bad;
        if (st) {
            st->dt = double(dt);
            st->iters_div = iv; st->iters_density = id;
            st->converged_div = cv; st->converged_density = cd;
            st->err_div = ev; st->err_density = ed;
            st->max_speed = double(vmax);
            st->nonfinite = bad;
        }
        return bad ? -5 : 0;
    }

    int get(int field, double* out) const override {
        auto cp3 = [&](const std::vector<R>& s) { for (size_t i = 0; i < s.size(); ++i) out[i] = double(s[i]); };
        switch (field) {
            case F_X: cp3(x); return 0;
            case F_V: cp3(v); return 0;
            case F_RHO: for (uint32_t i = 0; i < n; ++i) out[i] = double(da[4 * i]); return 0;
            case F_AOR2: for (uint32_t i = 0; i < n; ++i) out[i] = double(da[4 * i + 1]); return 0;
            case F_RHO_ADV: for (uint32_t i = 0; i < n; ++i) out[i] = double(da[4 * i + 2]); return 0;
            case F_RHO_CHANGE: for (uint32_t i = 0; i < n; ++i) out[i] = double(da[4 * i + 3]); return 0;
            case F_P_DIV: cp3(p_div_final); return 0;
            case F_P_DENSITY: cp3(p_den_final); return 0;
            case F_ACCEL: cp3(accel); return 0;
            case F_ALPHA: cp3(alpha); return 0;
            case F_BX: cp3(bx); return 0;
            case F_BV: cp3(bvol); return 0;
            case F_GF: cp3(gf); return 0;
            case F_GB: cp3(gb); return 0;
            case F_RES_DIV: case F_RES_DENSITY: {
                const std::vector<double>& r = field == F_RES_DIV ? res_div : res_den;
                for (uint32_t i = 0; i < n; ++i) out[i] = i < r.size() ? r[i] : 0.0;
                return 0;
            }
            default: return -1;
        }
    }
    int get_u32(int field, uint32_t* out) const override {
        const std::vector<uint32_t>* s = nullptr;
        switch (field) {
            case F_NBR: s = &nbr; break;
            case F_NBR_B: s = &nbr_b; break;
            case F_KEYS: s = &keys; break;
            case F_SORTED: s = &sorted; break;
            case F_CELL_STARTS: s = &starts; break;
            default: return -1;
        }
        std::memcpy(out, s->data(), s->size() * sizeof(uint32_t));
        return 0;
    }
    int set(int field, const double* in) override {
        switch (field) {
            case F_X: for (size_t i = 0; i < x.size(); ++i) x[i] = R(in[i]); return 0;
            case F_V: for (size_t i = 0; i < v.size(); ++i) v[i] = R(in[i]); return 0;
            case F_WARM_DIV: for (uint32_t i = 0; i < n; ++i) warm_div[i] = R(in[i]); return 0;
            case F_WARM_DENSITY: for (uint32_t i = 0; i < n; ++i) warm_den[i] = R(in[i]); return 0;
            default: return -1;
        }
    }
};

}  // namespace

extern "C" {

void oracle_derive(const oracle_config* c, oracle_derived* out) { *out = derive(*c); }

uint32_t oracle_morton3(uint32_t x, uint32_t y, uint32_t z) { return morton3(x, y, z); }

// W and |grad W| at (q, h) in the reference's float arithmetic: the norms are
// derived from h exactly as main.cpp:2012-2019 does.
void oracle_kernel_f32(float q, float h, float* W, float* gradW_mag) {
    const float pi = float(M_PI);
    float k = 8.0f / (pi * h * h * h);
    float g = 8.0f / (pi * h * h * h * h);
    *W = kernel_W(q, k);
    float r = q * h;
    *gradW_mag = (r > 0.0f) ? std::fabs(kernel_gradW_scale(r, q, g) * r) : 0.0f;
}
void oracle_kernel_f64(double q, double h, double* W, double* gradW_mag) {
    double k = 8.0 / (M_PI * h * h * h);
    double g = 8.0 / (M_PI * h * h * h * h);
    *W = kernel_W(q, k);
    double r = q * h;
    *gradW_mag = (r > 0.0) ? std::fabs(kernel_gradW_scale(r, q, g) * r) : 0.0;
}

// Initial fill of one preset: counts of seedInitialFill (main.cpp:2266-2294),
// positions of shaders/initial_fill.comp.glsl:22-65, float operations in the
// shader's order.  brick = {min xyz, max xyz}; droplet = {centre xyz, radius,
// velocity xyz}.  Returns the total count; writes positions, velocities and
// flags (count entries) when out_x != NULL.
uint32_t oracle_initial_fill(const float brick[6], int add_droplet, const float droplet[7], float r,
                             uint32_t capacity, float* out_x, float* out_v, uint32_t* out_flags,
                             uint32_t* out_brick) {
    const float sp = 2.0f * r;
    float ext[3] = {brick[3] - brick[0], brick[4] - brick[1], brick[5] - brick[2]};
    uint32_t n_axis[3];
    for (int a = 0; a < 3; ++a) {
        const uint32_t c = uint32_t(ext[a] / sp);           // main.cpp:2270-2272
        n_axis[a] = c > 1u ? c : 1u;
    }
    uint32_t nb = n_axis[0] * n_axis[1] * n_axis[2];
    if (nb > capacity) nb = capacity;                        // main.cpp:2273
    uint32_t nd = 0;
    if (add_droplet) {                                       // main.cpp:2277-2282
        uint32_t dr = uint32_t(droplet[3] / sp);
        if (dr < 1u) dr = 1u;
        nd = (4u * dr * dr * dr) / 3u;
        if (nd > capacity - nb) nd = capacity - nb;
    }
    if (out_brick) *out_brick = nb;
    const uint32_t total = nb + nd;
    if (!out_x) return total;
    // shader dims: uvec3(max(extent / spacing, 1.0)) (initial_fill.comp.glsl:31)
    uint32_t dims[3], dd[3];
    for (int a = 0; a < 3; ++a) {
        const float q = ext[a] / sp;
        dims[a] = uint32_t(q > 1.0f ? q : 1.0f);
        const float w = 2.0f * droplet[3] / sp;            // :44
        dd[a] = uint32_t(w > 1.0f ? w : 1.0f);
    }
    for (uint32_t g = 0; g < total; ++g) {
        float x[3], v[3] = {0.0f, 0.0f, 0.0f};
        uint32_t fl = 0u;
        const bool in_brick = g < nb;
        const uint32_t* d = in_brick ? dims : dd;
        const uint32_t t = in_brick ? g : g - nb;
        const uint32_t dx = d[0] ? d[0] : 1u, dxy = (d[0] * d[1]) != 0u ? d[0] * d[1] : 1u, dy = d[1] ? d[1] : 1u;
        const float c[3] = {float(t % dx), float((t / dx) % dy), float(t / dxy)};
        if (in_brick) {
            for (int a = 0; a < 3; ++a) x[a] = brick[a] + (c[a] + 0.5f) * sp;          // :41
        } else {
            float l[3];
            for (int a = 0; a < 3; ++a) l[a] = (c[a] - float(d[a]) * 0.5f + 0.5f) * sp;   // :48-49
            const float ll = std::sqrt(l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
            if (ll > droplet[3]) {                                                      // :51
                const float dn = ll > 1e-7f ? ll : 1e-7f;
                for (int a = 0; a < 3; ++a) l[a] = l[a] / dn * droplet[3] * 0.95f;
            }
            for (int a = 0; a < 3; ++a) {
                x[a] = droplet[a] + l[a];
                v[a] = droplet[4 + a];
            }
            fl = 2u;
        }
        for (int a = 0; a < 3; ++a) {
            out_x[3 * size_t(g) + a] = x[a];
            if (out_v) out_v[3 * size_t(g) + a] = v[a];
        }
        if (out_flags) out_flags[g] = fl;
    }
    return total;
}

// Emitter placement of shaders/apply_emitter.comp.glsl:29-65: particle g of a
// batch of k goes to slot n0 + g from emitter g % count; em holds count
// emitters as (pos xyz, radius, velocity xyz, rate).
static uint32_t pcg32_hash(uint32_t v) {                     // :29-33
    uint32_t st = v * 747796405u + 2891336453u;
    uint32_t w = ((st >> ((st >> 28u) + 4u)) ^ st) * 277803737u;
    return (w >> 22u) ^ w;
}
void oracle_emit(const float* em, uint32_t count, uint32_t n0, uint32_t k, float* out_x, float* out_v) {
    for (uint32_t g = 0; g < k; ++g) {
        const uint32_t e = g % count, slot = n0 + g;
        const uint32_t seed = slot * 7919u + e * 31u;             // :55
        float s[3];
        for (int a = 0; a < 3; ++a) {                            // :35-42
            const float u = float(pcg32_hash(seed + uint32_t(a))) / 4294967295.0f;
            s[a] = u * 2.0f - 1.0f;
        }
        const float len = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        if (len > 1.0f)
            for (int a = 0; a < 3; ++a) s[a] = s[a] / len;
        const float* E = em + 8 * size_t(e);
        for (int a = 0; a < 3; ++a) {
            out_x[3 * size_t(g) + a] = E[a] + s[a] * E[3];         // :56-57
            out_v[3 * size_t(g) + a] = E[4 + a];
        }
    }
}

// Boundary samples (main.cpp:1332-1431).  Returns the count; writes up to
// `cap_out` xyz triples when out != NULL.
uint32_t oracle_boundary_samples(const float dmin[3], const float dmax[3], float r, uint32_t cap,
                                 float* out, uint32_t cap_out) {
    std::vector<float> s = boundary_samples(dmin, dmax, r, cap);
    uint32_t cnt = uint32_t(s.size() / 3);
    if (out) std::memcpy(out, s.data(), sizeof(float) * 3 * std::min(cnt, cap_out));
    return cnt;
}

// Stand-alone hash over the reference grid (keys, sorted index, cell starts
// with sentinel).  `starts` must hold total_cells + 1 entries.
void oracle_hash(const oracle_config* c, uint32_t n, const float* x, uint32_t* keys,
                 uint32_t* sorted, uint32_t* starts) {
    oracle_derived d = derive(*c);
    std::vector<uint32_t> k, s, st;
    hash_grid<float>(n, x, c->dmin, d.cell, d.axes, d.total_cells, k, s, st);
    std::memcpy(keys, k.data(), n * sizeof(uint32_t));
    std::memcpy(sorted, s.data(), n * sizeof(uint32_t));
    std::memcpy(starts, st.data(), st.size() * sizeof(uint32_t));
}

void* oracle_create(const oracle_config* c, int fp64, uint32_t n, const float* x, const float* v) {
    if (!c || (n && !x)) return nullptr;
    if (fp64) return static_cast<SimBase*>(new Sim<double>(*c, n, x, v));
    return static_cast<SimBase*>(new Sim<float>(*c, n, x, v));
}
void oracle_destroy(void* s) { delete static_cast<SimBase*>(s); }
int oracle_step(void* s, float dt, const oracle_policy* p, oracle_stats* st) {
    return static_cast<SimBase*>(s)->step(dt, *p, st);
}
int oracle_get(void* s, int field, double* out) { return static_cast<SimBase*>(s)->get(field, out); }
int oracle_get_u32(void* s, int field, uint32_t* out) { return static_cast<SimBase*>(s)->get_u32(field, out); }
int oracle_set(void* s, int field, const double* in) { return static_cast<SimBase*>(s)->set(field, in); }
uint32_t oracle_count(void* s, int which) { return static_cast<SimBase*>(s)->count(which); }
void oracle_sim_derived(void* s, oracle_derived* out) { *out = static_cast<SimBase*>(s)->derived(); }
void oracle_prepare(void* s, int psi) { static_cast<SimBase*>(s)->prepare(psi); }
void oracle_apply_A(void* s, int mode_k, double dt, int psi, const double* p, double* ap, double* acc) {
    static_cast<SimBase*>(s)->apply_A(mode_k, dt, psi, p, ap, acc);
}
// Host threads of the oracle's parallel loops (the CPU baseline times 1 and all).
void oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int oracle_max_threads(void) { return omp_get_max_threads(); }
void oracle_source(void* s, int adv, double dt) { static_cast<SimBase*>(s)->source(adv, dt); }
void oracle_forces(void* s, double dt) { static_cast<SimBase*>(s)->forces(dt); }

}  // extern "C"
