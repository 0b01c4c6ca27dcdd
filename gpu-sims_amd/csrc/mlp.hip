// mlp.hip -- learned warm start for the DFSPH pressure solves.
//
// The reference zero-fills both pressure solves (main.cpp:2767-2769,
// :2823-2825).  Here a per-particle network 16 -> 64 -> 64 -> 64 (ReLU) -> 2
// predicts the converged stiffness of the divergence and density solves from
// 16 features; the solver then iterates to the same tolerance, so the plain
// solver stays the judge.  BF16 MFMA (v_mfma_f32_32x32x16_bf16) with FP32
// accumulation for every layer product and every backward product; biases,
// the 64 -> 2 output layer, the loss and Adam in FP32.
//
// Lane layout of a 32x32x16 tile (checked on gfx950): lane l supplies
// A[l % 32][8*(l/32) + t] and B[8*(l/32) + t][l % 32], t = 0..7, and receives
// C[8*(v/4) + 4*(l/32) + v%4][l % 32] in accumulator v.
#include "mlp.h"

#include "grid.h"

namespace dfsph {

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;
using bf16 = __bf16;

// Padded LDS row strides (elements): 24 for 16-wide rows, 72 for 64-wide rows,
// so the 32 rows a tile reads start in different banks.
constexpr int LX = 24;
constexpr int LH = 72;

// Packed BF16 weights (global, produced from the FP32 master by k_pack):
// forward operands W^T [n][k] and backward operands W [k][n].
struct WPack {
    bf16 w1t[kNH * LX];
    bf16 w2t[kNH * LH];
    bf16 w3t[kNH * LH];
    bf16 w2[kNH * LH];
    bf16 w3[kNH * LH];
};
constexpr int kFwdBytes = sizeof(bf16) * (kNH * LX + 2 * kNH * LH);
constexpr int kAllBytes = sizeof(WPack);

struct WSmall {   // FP32 parts, read from the master vector (16-byte multiple)
    float b1[kNH], b2[kNH], b3[kNH], w4[kNH * kNO], b4[kNO], pad[2];
};
static_assert(sizeof(WSmall) % 16 == 0, "LDS buffers after WSmall must stay 16-byte aligned");

__device__ __forceinline__ f32x16 zero16() {
    f32x16 z;
#pragma unroll
    for (int v = 0; v < 16; ++v) z[v] = 0.0f;
    return z;
}

// acc += A[32 x K] * B[K x 32] with A row-major (row stride lda) and B given by
// its transpose bt (row j = column j of B, stride ldb); 16-byte operand loads.
template <int K>
__device__ __forceinline__ f32x16 mma_nt(const bf16* a, int lda, const bf16* bt, int ldb, f32x16 acc) {
    const int l = threadIdx.x & 63, r = l & 31, kg = l >> 5;
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 16) {
        const bf16x8 av = *reinterpret_cast<const bf16x8*>(a + r * lda + k0 + kg * 8);
        const bf16x8 bv = *reinterpret_cast<const bf16x8*>(bt + r * ldb + k0 + kg * 8);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
    }
    return acc;
}

// acc += A[32 x 64] * B[64 x 32] with element strides (weight gradients, where
// the reduction runs over samples): A(i, s) = a[i*ai + s*as] for i < mrows,
// zero otherwise; B(s, j) = b[s*bs + j].
__device__ __forceinline__ f32x16 mma_samples(const bf16* a, int ai, int as, int mrows,
                                              const bf16* b, int bs, f32x16 acc) {
    const int l = threadIdx.x & 63, r = l & 31, kg = l >> 5;
#pragma unroll
    for (int k0 = 0; k0 < 64; k0 += 16) {
        bf16x8 av, bv;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int s = k0 + kg * 8 + t;
            av[t] = r < mrows ? a[r * ai + s * as] : bf16(0.0f);
            bv[t] = b[s * bs + r];
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
    }
    return acc;
}

// Row of accumulator v of this lane within a 32x32 tile.
__device__ __forceinline__ int acc_row(int v) { return 8 * (v >> 2) + 4 * ((threadIdx.x & 63) >> 5) + (v & 3); }

// One hidden layer for the wave's 64 samples: hout = relu(hin * W + b), K = kin.
template <int K>
__device__ __forceinline__ void layer(const bf16* hin, int ldin, const bf16* wt, int ldw,
                                      const float* bias, bf16* hout) {
    const int col = threadIdx.x & 31;
    f32x16 acc[2][2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
            acc[mt][nt] = mma_nt<K>(hin + mt * 32 * ldin, ldin, wt + nt * 32 * ldw, ldw, zero16());
    // All reads of hin are issued before the writes below (same wave, in order),
    // so hout may alias hin.
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = mt * 32 + acc_row(v), c = nt * 32 + col;
                hout[row * LH + c] = bf16(fmaxf(acc[mt][nt][v] + bias[c], 0.0f));
            }
}

// Cooperative copy of packed weights and FP32 parts into LDS.
__device__ __forceinline__ void load_weights(const MlpDev& m, bf16* wl, int bytes, WSmall* ws) {
    const uint4* src = reinterpret_cast<const uint4*>(m.wpack);
    uint4* dst = reinterpret_cast<uint4*>(wl);
    for (int q = threadIdx.x; q < bytes / 16; q += blockDim.x) dst[q] = src[q];
    for (int q = threadIdx.x; q < kNH; q += blockDim.x) {
        ws->b1[q] = m.theta[kOB1 + q];
        ws->b2[q] = m.theta[kOB2 + q];
        ws->b3[q] = m.theta[kOB3 + q];
        ws->w4[2 * q] = m.theta[kOW4 + 2 * q];
        ws->w4[2 * q + 1] = m.theta[kOW4 + 2 * q + 1];
    }
    if (threadIdx.x < kNO) ws->b4[threadIdx.x] = m.theta[kOB4 + threadIdx.x];
}

// Forward of the wave's 64 samples; x (64 x LX) holds the inputs, h (64 x LH)
// is scratch.  Returns this lane's two outputs (lane = sample).
__device__ __forceinline__ void forward(const bf16* wl, const WSmall* ws, const bf16* x, bf16* h1,
                                        bf16* h2, bf16* h3, float y[2]) {
    const WPack* w = reinterpret_cast<const WPack*>(wl);
    layer<16>(x, LX, w->w1t, LX, ws->b1, h1);
    layer<64>(h1, LH, w->w2t, LH, ws->b2, h2);
    layer<64>(h2, LH, w->w3t, LH, ws->b3, h3);
    const int s = threadIdx.x & 63;
    float y0 = ws->b4[0], y1 = ws->b4[1];
#pragma unroll 8
    for (int k = 0; k < kNH; ++k) {
        const float hv = float(h3[s * LH + k]);
        y0 += hv * ws->w4[2 * k];
        y1 += hv * ws->w4[2 * k + 1];
    }
    y[0] = y0;
    y[1] = y1;
}

// ---- features -------------------------------------------------------------
// 16 per-particle inputs for solve `phase` (0 divergence, 1 density), all
// dimensionless; heavy-tailed quantities go through slog(x) = sign(x) log(1+|x|):
//  0 density deviation slog((rho/rho0 - 1)*100)    9-11 position (x - dmin)/(dmax - dmin)
//  1 divergence source slog(-rho_dot*dt*1e3)        12 boundary neighbours / 30
//  2 log(alpha*rho^3)/10 (inverse Jacobi diagonal)  13 |v| dt/(2r)
//  3 log(1 + previous kappa_v)                      14 density source slog((1 - rho*)*1e3), phase 1
//  4 log(1 + previous kappa_rho)                    15 phase
//  5 fluid neighbours / 30
//  6-8 velocity v dt/(2r)
// kappa is the solve's pressure in units of its inverse Jacobi diagonal:
// kappa_v = p_div dt/(alpha rho^3) 1e3, kappa_rho = p dt^2/(alpha rho^3) 1e3.
// The network's outputs are log(1 + kappa).
__device__ __forceinline__ float slog(float x) { return copysignf(log1pf(fabsf(x)), x); }

__device__ __forceinline__ void features(const SimConst& k, const Buffers& b, uint32_t i, int phase,
                                         float dt, float dt_prev, float f[kNF]) {
    const float rho = b.rho[i];
    const float cbase = b.alpha[i] * rho * rho * rho;
    const float ratio = dt_prev > 0.0f ? dt_prev / dt : 0.0f;
    const float4 x = static_cast<const float4*>(b.pos4)[i];
    const float4 v = static_cast<const float4*>(phase ? b.vel4b : b.vel4)[i];
    const float vs = dt / (0.5f * k.h);   // dt / (2r)
    const float inv_c = cbase > 0.0f ? 1.0f / cbase : 0.0f;
    f[0] = slog((rho / k.rho0 - 1.0f) * 100.0f);
    f[1] = slog(-b.rho_dot[i] * dt * 1000.0f);
    f[2] = logf(fmaxf(cbase, 1e-30f)) * 0.1f;
    f[3] = log1pf(fmaxf(b.pdiv_prev[i] * ratio * dt * inv_c * 1000.0f, 0.0f));
    f[4] = log1pf(fmaxf(b.pden_prev[i] * ratio * ratio * dt * dt * inv_c * 1000.0f, 0.0f));
    f[5] = float(b.nbr[i]) * (1.0f / 30.0f);
    f[6] = v.x * vs;
    f[7] = v.y * vs;
    f[8] = v.z * vs;
    f[9] = (x.x - k.dmin[0]) / (k.dmax[0] - k.dmin[0]);
    f[10] = (x.y - k.dmin[1]) / (k.dmax[1] - k.dmin[1]);
    f[11] = (x.z - k.dmin[2]) / (k.dmax[2] - k.dmin[2]);
    f[12] = float(b.nbrb[i]) * (1.0f / 30.0f);
    f[13] = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z) * vs;
    f[14] = phase ? slog((1.0f - b.rho_adv[i]) * 1000.0f) : 0.0f;
    f[15] = float(phase);
#pragma unroll
    for (int q = 0; q < kNF; ++q)
        if (!isfinite(f[q])) f[q] = 0.0f;
}

// Network output -> kappa, clamped to [0, e^12 - 1].
__device__ __forceinline__ float kappa_of(float y) { return expm1f(fminf(fmaxf(y, 0.0f), 12.0f)); }

// ---- inference in the step: features -> network -> initial pressure --------
constexpr int kInferWaves = 4;

__global__ void __launch_bounds__(64 * kInferWaves) k_mlp_warm(SimConst k, Buffers b, MlpDev m, int phase) {
    extern __shared__ __align__(16) unsigned char smem[];
    bf16* wl = reinterpret_cast<bf16*>(smem);
    WSmall* ws = reinterpret_cast<WSmall*>(smem + kFwdBytes);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    bf16* xb = reinterpret_cast<bf16*>(smem + kFwdBytes + sizeof(WSmall)) + wave * (64 * LX + 64 * LH);
    bf16* hb = xb + 64 * LX;
    load_weights(m, wl, kFwdBytes, ws);
    __syncthreads();
    const float dt = b.status->dt, dt_prev = b.status->dt_prev;
    const uint32_t i = (blockIdx.x * kInferWaves + wave) * 64 + lane;
    float f[kNF];
    if (i < k.n) features(k, b, i, phase, dt, dt_prev, f);
    else
#pragma unroll
        for (int q = 0; q < kNF; ++q) f[q] = 0.0f;
#pragma unroll
    for (int q = 0; q < kNF; ++q) xb[lane * LX + q] = bf16(f[q]);
    float y[2];
    forward(wl, ws, xb, hb, hb, hb, y);
    if (i < k.n) {
        // The guess is kappa in units of the inverse Jacobi diagonal; for a
        // particle with almost no neighbours alpha (and so that unit) is up to
        // 13 orders larger than in the bulk, where the network's small output
        // error would become a huge pressure.  The unit is taken with alpha
        // capped at 10x a full neighbourhood's, so the guess stays bounded
        // there; the solve brings such a particle's pressure to its value.
        const float rho = b.rho[i];
        const float cbase = fminf(b.alpha[i], k.alpha_cap) * rho * rho * rho;
        const float c = phase ? cbase / (dt * dt) : cbase / dt;
        const float p0 = kappa_of(y[phase]) * 1e-3f * c;
        const float p = isfinite(p0) ? p0 : 0.0f;
        b.p[i] = p;
        static_cast<float4*>(b.pos4)[i].w = p / fmaxf(rho * rho, 1.0e-5f);   // p / rho^2
    }
}

// ---- dataset recording ---------------------------------------------------------
// A particle is sampled in a step when (id + 7919*step) % stride == 0; its two
// rows (one per phase) live at base + 2*(id / stride) + phase.
__device__ __forceinline__ bool sampled(uint32_t id, uint32_t step, uint32_t stride) {
    return (id + 7919u * step) % stride == 0u;
}

__global__ void __launch_bounds__(256) k_record_features(SimConst k, Buffers b, MlpDev m, int phase,
                                                         uint32_t step, uint32_t stride, uint32_t base,
                                                         uint32_t cap) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k.n) return;
    const uint32_t id = b.ids[i];
    if (!sampled(id, step, stride)) return;
    const uint32_t row = base + 2u * (id / stride) + uint32_t(phase);
    if (row >= cap) return;
    float f[kNF];
    features(k, b, i, phase, b.status->dt, b.status->dt_prev, f);
#pragma unroll
    for (int q = 0; q < kNF; ++q) m.ds_x[size_t(row) * kNF + q] = f[q];
}

// Targets after both solves: the converged kappa_v and kappa_rho, for both rows.
__global__ void __launch_bounds__(256) k_record_targets(SimConst k, Buffers b, MlpDev m, uint32_t step,
                                                        uint32_t stride, uint32_t base, uint32_t cap) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k.n) return;
    const uint32_t id = b.ids[i];
    if (!sampled(id, step, stride)) return;
    const uint32_t row = base + 2u * (id / stride);
    const float dt = b.status->dt;
    const float rho = b.rho[i];
    const float cbase = b.alpha[i] * rho * rho * rho;
    const float inv_c = cbase > 0.0f ? 1.0f / cbase : 0.0f;
    const float kv = log1pf(fmaxf(b.p_div[i] * dt * inv_c * 1000.0f, 0.0f));
    const float kd = log1pf(fmaxf(b.p_den[i] * dt * dt * inv_c * 1000.0f, 0.0f));
    for (uint32_t r = row; r < row + 2u && r < cap; ++r) {
        m.ds_t[size_t(r) * kNO] = kv;
        m.ds_t[size_t(r) * kNO + 1] = kd;
    }
}

// ---- training: one wave = 64 samples; per-workgroup gradient partials ---------
constexpr int kTrainSmem = kAllBytes + int(sizeof(WSmall)) + 2 * (64 * LX + 4 * 64 * LH) + 64 * 2 * 4;

__global__ void __launch_bounds__(64) k_mlp_train(MlpDev m, const uint32_t* __restrict__ rows,
                                                  uint32_t nrows, float inv_batch, float* __restrict__ part,
                                                  float* __restrict__ part_loss) {
    extern __shared__ __align__(16) unsigned char smem[];
    bf16* wl = reinterpret_cast<bf16*>(smem);
    const WPack* w = reinterpret_cast<const WPack*>(wl);
    WSmall* ws = reinterpret_cast<WSmall*>(smem + kAllBytes);
    bf16* x = reinterpret_cast<bf16*>(smem + kAllBytes + sizeof(WSmall));
    bf16* h1 = x + 64 * LX;
    bf16* h2 = h1 + 64 * LH;
    bf16* h3 = h2 + 64 * LH;
    bf16* dh = h3 + 64 * LH;
    float* dy = reinterpret_cast<float*>(dh + 64 * LH);
    load_weights(m, wl, kAllBytes, ws);
    __syncthreads();
    const int s = threadIdx.x;                       // sample slot = lane
    const uint32_t q = blockIdx.x * 64u + uint32_t(s);
    const bool valid = q < nrows;
    const uint32_t row = valid ? rows[q] : 0u;
#pragma unroll
    for (int c = 0; c < kNF; ++c) x[s * LX + c] = bf16(valid ? m.ds_x[size_t(row) * kNF + c] : 0.0f);
    float y[2];
    forward(wl, ws, x, h1, h2, h3, y);
    // MSE over the two outputs, mean over the batch.
    float d0 = 0.0f, d1 = 0.0f;
    if (valid) {
        d0 = y[0] - m.ds_t[size_t(row) * kNO];
        d1 = y[1] - m.ds_t[size_t(row) * kNO + 1];
    }
    float l2 = d0 * d0 + d1 * d1;
    dy[2 * s] = 2.0f * d0 * inv_batch;
    dy[2 * s + 1] = 2.0f * d1 * inv_batch;
    // dH3 = (dY W4^T) * [h3 > 0]
#pragma unroll 8
    for (int kk = 0; kk < kNH; ++kk) {
        const float g = dy[2 * s] * ws->w4[2 * kk] + dy[2 * s + 1] * ws->w4[2 * kk + 1];
        dh[s * LH + kk] = bf16(float(h3[s * LH + kk]) > 0.0f ? g : 0.0f);
    }
    float* P = part + size_t(blockIdx.x) * kNParam;
    // dW4[k][o] = sum_s h3[s][k] dy[s][o]; db4; db3[n] = sum_s dh3[s][n]  (lane = unit)
    {
        const int u = threadIdx.x;
        float g0 = 0.0f, g1 = 0.0f, gb = 0.0f;
        for (int t = 0; t < 64; ++t) {
            const float hv = float(h3[t * LH + u]);
            g0 += hv * dy[2 * t];
            g1 += hv * dy[2 * t + 1];
            gb += float(dh[t * LH + u]);
        }
        P[kOW4 + 2 * u] = g0;
        P[kOW4 + 2 * u + 1] = g1;
        P[kOB3 + u] = gb;
        if (u < kNO) {
            float b4 = 0.0f;
            for (int t = 0; t < 64; ++t) b4 += dy[2 * t + u];
            P[kOB4 + u] = b4;
        }
    }
    const int col = threadIdx.x & 31;
    // Weight gradient of a 64-wide layer: dW[k][n] = sum_s hin[s][k] dh[s][n].
    auto wgrad = [&](const bf16* hin, int ldin, int kin, int off) {
        for (int mt = 0; mt * 32 < kin; ++mt)
            for (int nt = 0; nt < 2; ++nt) {
                f32x16 acc = mma_samples(hin + mt * 32, 1, ldin, kin - mt * 32, dh + nt * 32, LH, zero16());
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int kk = mt * 32 + acc_row(v);
                    if (kk < kin) P[off + kk * kNH + nt * 32 + col] = acc[v];
                }
            }
    };
    // Back-propagate through a layer: dh <- (dh W^T) * [hprev > 0]; W given as w[k][n].
    auto backprop = [&](const bf16* wkn, const bf16* hprev) {
        f32x16 acc[2][2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int nt = 0; nt < 2; ++nt)
                acc[mt][nt] = mma_nt<64>(dh + mt * 32 * LH, LH, wkn + nt * 32 * LH, LH, zero16());
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int nt = 0; nt < 2; ++nt)
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int r = mt * 32 + acc_row(v), c = nt * 32 + col;
                    dh[r * LH + c] = bf16(float(hprev[r * LH + c]) > 0.0f ? acc[mt][nt][v] : 0.0f);
                }
    };
    auto bgrad = [&](int off) {
        float gb = 0.0f;
        for (int t = 0; t < 64; ++t) gb += float(dh[t * LH + threadIdx.x]);
        P[off + threadIdx.x] = gb;
    };
    wgrad(h2, LH, kNH, kOW3);
    backprop(w->w3, h2);        // dh = dH2
    bgrad(kOB2);
    wgrad(h1, LH, kNH, kOW2);
    backprop(w->w2, h1);        // dh = dH1
    bgrad(kOB1);
    wgrad(x, LX, kNF, kOW1);
    // Loss partial (wave sum).
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) l2 += __shfl_down(l2, o, 64);
    if (threadIdx.x == 0) part_loss[blockIdx.x] = l2;
}

// Fixed-order sum of the partials, then Adam on the FP32 master (or just the
// summed gradient when lr == 0).
__global__ void __launch_bounds__(256) k_mlp_adam(MlpDev m, const float* __restrict__ part, uint32_t nparts,
                                                  float lr, float b1, float b2, float eps, float c1, float c2,
                                                  float* __restrict__ grad_out) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= uint32_t(kNParam)) return;
    float g = 0.0f;
    for (uint32_t w = 0; w < nparts; ++w) g += part[size_t(w) * kNParam + p];
    if (grad_out) grad_out[p] = g;
    if (lr == 0.0f) return;
    const float mm = b1 * m.adam_m[p] + (1.0f - b1) * g;
    const float vv = b2 * m.adam_v[p] + (1.0f - b2) * g * g;
    m.adam_m[p] = mm;
    m.adam_v[p] = vv;
    m.theta[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// FP32 master -> packed BF16 operands (zero padding included).
__global__ void __launch_bounds__(256) k_mlp_pack(MlpDev m) {
    WPack* w = reinterpret_cast<WPack*>(m.wpack);
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < uint32_t(kNH * LX)) {
        const int n = t / LX, kk = t % LX;
        w->w1t[t] = bf16(kk < kNF ? m.theta[kOW1 + kk * kNH + n] : 0.0f);
    }
    if (t < uint32_t(kNH * LH)) {
        const int a = t / LH, c = t % LH;
        const bool in = c < kNH;
        w->w2t[t] = bf16(in ? m.theta[kOW2 + c * kNH + a] : 0.0f);   // [n=a][k=c]
        w->w3t[t] = bf16(in ? m.theta[kOW3 + c * kNH + a] : 0.0f);
        w->w2[t] = bf16(in ? m.theta[kOW2 + a * kNH + c] : 0.0f);    // [k=a][n=c]
        w->w3[t] = bf16(in ? m.theta[kOW3 + a * kNH + c] : 0.0f);
    }
}

// Forward on caller-supplied rows (evaluation and tests): out[r][0..1].
__global__ void __launch_bounds__(64) k_mlp_eval(MlpDev m, const float* __restrict__ xin, uint32_t n,
                                                 float* __restrict__ out) {
    extern __shared__ __align__(16) unsigned char smem[];
    bf16* wl = reinterpret_cast<bf16*>(smem);
    WSmall* ws = reinterpret_cast<WSmall*>(smem + kFwdBytes);
    bf16* x = reinterpret_cast<bf16*>(smem + kFwdBytes + sizeof(WSmall));
    bf16* h = x + 64 * LX;
    load_weights(m, wl, kFwdBytes, ws);
    __syncthreads();
    const uint32_t r = blockIdx.x * 64u + threadIdx.x;
#pragma unroll
    for (int c = 0; c < kNF; ++c) x[threadIdx.x * LX + c] = bf16(r < n ? xin[size_t(r) * kNF + c] : 0.0f);
    float y[2];
    forward(wl, ws, x, h, h, h, y);
    if (r < n) {
        out[2 * size_t(r)] = y[0];
        out[2 * size_t(r) + 1] = y[1];
    }
}

inline uint32_t blocks_for(uint64_t n, uint32_t t) { return uint32_t((n + t - 1) / t); }
inline hipStream_t S(void* s) { return static_cast<hipStream_t>(s); }

constexpr int kInferSmem = kFwdBytes + int(sizeof(WSmall)) + kInferWaves * 2 * (64 * LX + 64 * LH);
constexpr int kEvalSmem = kFwdBytes + int(sizeof(WSmall)) + 2 * (64 * LX + 64 * LH);

}  // namespace

size_t mlp_pack_bytes() { return sizeof(WPack); }

void launch_mlp_warm(const SimConst& k, const Buffers& b, const MlpDev& m, int phase, void* stream) {
    if (!k.n) return;
    hipLaunchKernelGGL(k_mlp_warm, dim3(blocks_for(k.n, 64 * kInferWaves)), dim3(64 * kInferWaves),
                       kInferSmem, S(stream), k, b, m, phase);
}
void launch_record_features(const SimConst& k, const Buffers& b, const MlpDev& m, int phase,
                            uint32_t step, uint32_t stride, uint32_t base, uint32_t cap, void* stream) {
    if (!k.n) return;
    hipLaunchKernelGGL(k_record_features, dim3(blocks_for(k.n, 256)), dim3(256), 0, S(stream), k, b, m,
                       phase, step, stride, base, cap);
}
void launch_record_targets(const SimConst& k, const Buffers& b, const MlpDev& m, uint32_t step,
                           uint32_t stride, uint32_t base, uint32_t cap, void* stream) {
    if (!k.n) return;
    hipLaunchKernelGGL(k_record_targets, dim3(blocks_for(k.n, 256)), dim3(256), 0, S(stream), k, b, m,
                       step, stride, base, cap);
}
void launch_mlp_pack(const MlpDev& m, void* stream) {
    hipLaunchKernelGGL(k_mlp_pack, dim3(blocks_for(kNH * LH, 256)), dim3(256), 0, S(stream), m);
}
void launch_mlp_train(const MlpDev& m, const uint32_t* rows, uint32_t nrows, float* part, float* part_loss,
                      void* stream) {
    hipLaunchKernelGGL(k_mlp_train, dim3(blocks_for(nrows, 64)), dim3(64), kTrainSmem, S(stream), m, rows,
                       nrows, 1.0f / float(nrows), part, part_loss);
}
void launch_mlp_adam(const MlpDev& m, const float* part, uint32_t nparts, float lr, float b1, float b2,
                     float eps, float c1, float c2, float* grad_out, void* stream) {
    hipLaunchKernelGGL(k_mlp_adam, dim3(blocks_for(kNParam, 256)), dim3(256), 0, S(stream), m, part, nparts,
                       lr, b1, b2, eps, c1, c2, grad_out);
}
void launch_mlp_eval(const MlpDev& m, const float* x, uint32_t n, float* out, void* stream) {
    if (!n) return;
    hipLaunchKernelGGL(k_mlp_eval, dim3(blocks_for(n, 64)), dim3(64), kEvalSmem, S(stream), m, x, n, out);
}
int mlp_set_smem_limits() {
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(k_mlp_warm),
                            hipFuncAttributeMaxDynamicSharedMemorySize, kInferSmem) != hipSuccess)
        return 0;
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(k_mlp_train),
                            hipFuncAttributeMaxDynamicSharedMemorySize, kTrainSmem) != hipSuccess)
        return 0;
    return 1;
}
int mlp_smem_ok() {
    return kTrainSmem <= 160 * 1024 && kInferSmem <= 160 * 1024 && kEvalSmem <= 64 * 1024 ? 1 : 0;
}

}  // namespace dfsph
