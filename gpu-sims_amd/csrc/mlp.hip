// mlp.hip -- learned warm start for the DFSPH pressure solves.
//
// The reference zero-fills both pressure solves (main.cpp:2767-2769,
// :2823-2825).  Here a per-particle network 16 -> 64 -> 64 -> 64 (ReLU) -> 2
// predicts the converged stiffness of the divergence and density solves from
// 16 features; the solver then iterates to the same tolerance, so the plain
// solver still sets the accuracy.  BF16 MFMA (v_mfma_f32_32x32x16_bf16) with FP32
// accumulation for every layer product and every backward product; biases,
// the loss and Adam in FP32.  The 64 -> 2 output layer is FP32 in training and
// two BF16 products (W4's BF16 part and its BF16 rest) in inference.
//
// Lane layout of a 32x32x16 tile (checked on gfx950): lane l supplies
// A[l % 32][8*(l/32) + t] and B[8*(l/32) + t][l % 32], t = 0..7, and receives
// C[8*(v/4) + 4*(l/32) + v%4][l % 32] in accumulator v.
#include "mlp.h"

#include <algorithm>

#include "grid.h"
#include "launch.h"

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
constexpr int kAllBytes = sizeof(WPack);

struct WSmall {   // FP32 parts, read from the master vector (16-byte multiple)
    float b1[kNH], b2[kNH], b3[kNH], w4[kNH * kNO], b4[kNO], pad[2];
};
static_assert(sizeof(WSmall) % 16 == 0, "LDS buffers after WSmall must stay 16-byte aligned");

// Register operands of inference, in lane order: chunk c, lane l, 8 BF16 each
// (one 16-byte load per chunk and lane).  Chunks 0-1: layer 1 (output tile
// 0, 1); 2-9: layer 2 (tile * 4 + K block); 10-17: layer 3; 18-21: output
// layer (K block), rows 0-1 the BF16 part of W4 and rows 2-3 the BF16 rest.
constexpr int kRegChunks = 22;
struct WPackAll {
    WPack lds;
    bf16 lreg[kRegChunks * 64 * 8];
};
static_assert(sizeof(WPack) % 16 == 0, "register operands must stay 16-byte aligned");

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
// The network's outputs are log(1 + kappa).  Features are network inputs, not
// results: they use the hardware logarithm and reciprocal (about 1 ulp), and
// recording and inference compute them with the same code.
__device__ __forceinline__ float flog(float x) { return __logf(x); }
__device__ __forceinline__ float slog(float x) { return copysignf(flog(1.0f + fabsf(x)), x); }

// Per-launch constants of the features.
struct FeatK {
    float inv_rho0, vs, dt, ratio, dt_k1, dt_k2;   // dt_k1 = ratio dt 1e3, dt_k2 = (ratio dt)^2 1e3
    float dmin[3], inv_ext[3];
};

__device__ __forceinline__ FeatK feat_k(const SimConst& k, float dt, float dt_prev) {
    FeatK c;
    c.inv_rho0 = 1.0f / k.rho0;
    c.vs = dt / (0.5f * k.h);   // dt / (2r)
    c.dt = dt;
    c.ratio = dt_prev > 0.0f ? dt_prev / dt : 0.0f;
    c.dt_k1 = c.ratio * dt * 1000.0f;
    c.dt_k2 = c.ratio * c.ratio * dt * dt * 1000.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        c.dmin[a] = k.dmin[a];
        c.inv_ext[a] = 1.0f / (k.dmax[a] - k.dmin[a]);
    }
    return c;
}

// Raw per-particle inputs of the features (loaded separately so inference can
// fetch the next tile while the network runs on the current one).
struct FeatIn {
    float4 x, v;
    float rho, alpha, rho_dot, pdp, pdn, rho_adv;
    uint32_t nbr, nbrb;
};

__device__ __forceinline__ FeatIn load_feat_in(const Buffers& b, uint32_t i, int phase) {
    FeatIn r;
    r.rho = b.rho[i];
    r.alpha = b.alpha[i];
    r.x = static_cast<const float4*>(b.pos4)[i];
    r.v = static_cast<const float4*>(phase ? b.vel4b : b.vel4)[i];
    r.rho_dot = b.rho_dot[i];
    r.pdp = b.pdiv_prev[i];
    r.pdn = b.pden_prev[i];
    r.nbr = b.nbr[i];
    r.nbrb = b.nbrb[i];
    r.rho_adv = phase ? b.rho_adv[i] : 1.0f;
    return r;
}

__device__ __forceinline__ void features_from(const FeatK& c, const FeatIn& in, int phase, float f[kNF]) {
    const float rho = in.rho;
    const float cbase = in.alpha * rho * rho * rho;
    const float4 x = in.x;
    const float4 v = in.v;
    const float inv_c = cbase > 0.0f ? __frcp_rn(cbase) : 0.0f;
    f[0] = slog((rho * c.inv_rho0 - 1.0f) * 100.0f);
    f[1] = slog(-in.rho_dot * c.dt * 1000.0f);
    f[2] = flog(fmaxf(cbase, 1e-30f)) * 0.1f;
    f[3] = flog(1.0f + fmaxf(in.pdp * c.dt_k1 * inv_c, 0.0f));
    f[4] = flog(1.0f + fmaxf(in.pdn * c.dt_k2 * inv_c, 0.0f));
    f[5] = float(in.nbr) * (1.0f / 30.0f);
    f[6] = v.x * c.vs;
    f[7] = v.y * c.vs;
    f[8] = v.z * c.vs;
    f[9] = (x.x - c.dmin[0]) * c.inv_ext[0];
    f[10] = (x.y - c.dmin[1]) * c.inv_ext[1];
    f[11] = (x.z - c.dmin[2]) * c.inv_ext[2];
    f[12] = float(in.nbrb) * (1.0f / 30.0f);
    f[13] = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z) * c.vs;
    f[14] = phase ? slog((1.0f - in.rho_adv) * 1000.0f) : 0.0f;
    f[15] = float(phase);
#pragma unroll
    for (int q = 0; q < kNF; ++q)
        if (!isfinite(f[q])) f[q] = 0.0f;
}

// Network output -> kappa, clamped to [0, e^12 - 1].
__device__ __forceinline__ float kappa_of(float y) { return expm1f(fminf(fmaxf(y, 0.0f), 12.0f)); }

// ---- register-resident inference -------------------------------------------
// The network lives in VGPRs as MFMA A operands (output unit x layer input)
// and the activations are the B operand (layer input x 32 particles), so a
// layer's accumulators feed the next layer without leaving the lane: lane
// (r = l % 32, h = l / 32) receives rows 8g + 4h + j (g, j = 0..3) of each
// 32-row output tile, and K block kb of the next product takes, in slot t,
// row perm_k(kb, h, t) -- accumulators 8 (kb % 2) + t of tile kb / 2.  A
// product may order its reduction index freely; the packed weights follow
// the same order.  Biases enter as the accumulator input of the first product.
__device__ __host__ __forceinline__ int perm_k(int kb, int h, int t) {
    return 32 * (kb >> 1) + 8 * (2 * (kb & 1) + (t >> 2)) + 4 * h + (t & 3);
}

// Value of slot t of chunk c for lane l (see WPackAll).
__device__ __forceinline__ float lreg_value(const float* __restrict__ theta, int c, int l, int t) {
    const int r = l & 31, h = l >> 5;
    if (c < 2) return theta[kOW1 + (8 * h + t) * kNH + 32 * c + r];
    if (c < 18) {
        const int q = (c - 2) & 7, off = c < 10 ? kOW2 : kOW3;
        return theta[off + perm_k(q & 3, h, t) * kNH + 32 * (q >> 2) + r];
    }
    if (r >= 2 * kNO) return 0.0f;
    const float w = theta[kOW4 + perm_k(c - 18, h, t) * kNO + (r & 1)];
    const float hi = float(bf16(w));
    return r < kNO ? hi : w - hi;      // W4 = hi + rest, both exact in BF16 to 2^-17
}

// Inference operands in LDS: the lane-ordered chunks, then b1, b2, b3.
struct InferLds {
    bf16x8 w[kRegChunks * 64];
    float bl[3 * kNH];
};

// Cooperative copy of the register operands and the hidden biases into LDS;
// the caller synchronises.
__device__ __forceinline__ void load_infer(const MlpDev& m, InferLds* L) {
    const bf16x8* src = reinterpret_cast<const bf16x8*>(static_cast<const WPackAll*>(m.wpack)->lreg);
    for (int q = threadIdx.x; q < kRegChunks * 64; q += blockDim.x) L->w[q] = src[q];
    for (int q = threadIdx.x; q < 3 * kNH; q += blockDim.x) {
        const int layer = q / kNH, u = q % kNH;
        L->bl[q] = m.theta[(layer == 0 ? kOB1 : layer == 1 ? kOB2 : kOB3) + u];
    }
}

// Bias rows of output tile mo for this lane's accumulators (bl in LDS).
__device__ __forceinline__ f32x16 bias_tile(const float* bl, int mo, int h) {
    f32x16 c;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const float4 q = *reinterpret_cast<const float4*>(bl + 32 * mo + 8 * g + 4 * h);
        c[4 * g] = q.x;
        c[4 * g + 1] = q.y;
        c[4 * g + 2] = q.z;
        c[4 * g + 3] = q.w;
    }
    return c;
}

// BF16 (round to nearest even), then ReLU: accumulators base .. base + 7 as
// one B operand.  Rounding keeps the sign, so ReLU after rounding equals
// rounding after ReLU; on the BF16 bits it is a signed 16-bit max with 0.
typedef __attribute__((ext_vector_type(4))) uint32_t u32x4;

typedef __attribute__((ext_vector_type(2))) float f32x2;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;
typedef __attribute__((ext_vector_type(2))) short s16x2;

// One packed conversion (v_cvt_pk_bf16_f32) and one packed max per pair.
__device__ __forceinline__ uint32_t relu_pair(float x, float y) {
    const s16x2 q = __builtin_bit_cast(s16x2, __builtin_convertvector((f32x2{x, y}), bf16x2));
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(q, s16x2{0, 0}));
}

__device__ __forceinline__ bf16x8 relu_bf16(const f32x16& a, int base) {
    const u32x4 u{relu_pair(a[base], a[base + 1]), relu_pair(a[base + 2], a[base + 3]),
                  relu_pair(a[base + 4], a[base + 5]), relu_pair(a[base + 6], a[base + 7])};
    return __builtin_bit_cast(bf16x8, u);
}

// Forward of a wave's 64 particles: lane l holds the 16 features of particle
// l and receives its two outputs.  L holds the operands (load_infer).
__device__ __forceinline__ void forward_regs(const InferLds* L, const float f[kNF], float b40, float b41,
                                             float& y0, float& y1) {
    const int lane = threadIdx.x & 63, h = lane >> 5;
    const bf16x8* w = L->w + lane;                // chunk c at w[64 c]
    // Layer-1 B operand of N tile nt, lane (n, h): features 8h .. 8h+7 of
    // particle 32 nt + n.  Lane l has particle l: its own half for nt = h,
    // the partner lane's (l ^ 32) for the other tile.
    bf16x8 own, snd;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        own[t] = bf16(h ? f[8 + t] : f[t]);
        snd[t] = bf16(h ? f[t] : f[8 + t]);
    }
    const u32x4 s4 = __builtin_bit_cast(u32x4, snd);
    u32x4 r4;
#pragma unroll
    for (int q = 0; q < 4; ++q) r4[q] = __shfl_xor(s4[q], 32, 64);
    const bf16x8 rcv = __builtin_bit_cast(bf16x8, r4);
    bf16x8 x[2];
    x[0] = h ? rcv : own;
    x[1] = h ? own : rcv;
    f32x16 acc[2][2];                              // [output tile][N tile]
#pragma unroll
    for (int mo = 0; mo < 2; ++mo) {
        const f32x16 c = bias_tile(L->bl, mo, h);
        const bf16x8 a = w[64 * mo];
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) acc[mo][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, x[nt], c, 0, 0, 0);
    }
    auto hidden = [&](int c0, const float* b) {
        bf16x8 in[2][4];
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) in[nt][kb] = relu_bf16(acc[kb >> 1][nt], 8 * (kb & 1));
#pragma unroll
        for (int mo = 0; mo < 2; ++mo) {
            const f32x16 c = bias_tile(b, mo, h);
            f32x16 a0 = c, a1 = c;
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) {
                const bf16x8 wa = w[64 * (c0 + 4 * mo + kb)];
                a0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa, in[0][kb], a0, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa, in[1][kb], a1, 0, 0, 0);
            }
            acc[mo][0] = a0;
            acc[mo][1] = a1;
        }
    };
    hidden(2, L->bl + kNH);
    hidden(10, L->bl + 2 * kNH);
    // Output layer: rows 0-1 (W4's BF16 part) and 2-3 (its rest) of one tile;
    // lanes h = 0 hold them.
    f32x16 o0 = zero16(), o1 = zero16();
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
        const bf16x8 wa = w[64 * (18 + kb)];
        o0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa, relu_bf16(acc[kb >> 1][0], 8 * (kb & 1)), o0, 0, 0, 0);
        o1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa, relu_bf16(acc[kb >> 1][1], 8 * (kb & 1)), o1, 0, 0, 0);
    }
    const float s0 = __shfl_xor(o1[0] + o1[2], 32, 64), s1 = __shfl_xor(o1[1] + o1[3], 32, 64);
    y0 = (h ? s0 : o0[0] + o0[2]) + b40;
    y1 = (h ? s1 : o0[1] + o0[3]) + b41;
}

// ---- inference in the step: features -> network -> initial pressure --------
// Persistent waves: each loads the network once and walks tiles of 64
// particles, fetching the next tile's inputs before running the current one.
constexpr int kInferThreads = 256;

__global__ void __launch_bounds__(kInferThreads) k_mlp_warm(SimConst k, Buffers b, MlpDev m, int phase) {
    __shared__ InferLds L;
    load_infer(m, &L);
    const float b40 = m.theta[kOB4], b41 = m.theta[kOB4 + 1];
    __syncthreads();
    const float dt = b.status->dt;
    const FeatK fk = feat_k(k, dt, b.status->dt_prev);
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t waves = gridDim.x * (blockDim.x / 64);
    const uint32_t ntiles = (k.n + 63u) / 64u;
    uint32_t t = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);   // wave-uniform
    FeatIn cur{};
    if (t < ntiles && t * 64u + lane < k.n) cur = load_feat_in(b, t * 64u + lane, phase);
    for (; t < ntiles; t += waves) {
        // Compiler-only memory barrier: keeps the per-tile operand reads from
        // LDS inside the loop (hoisted, they would hold ~180 VGPRs).
        asm volatile("" ::: "memory");
        const uint32_t i = t * 64u + lane, tn = t + waves;
        FeatIn nxt{};
        if (tn < ntiles && tn * 64u + lane < k.n) nxt = load_feat_in(b, tn * 64u + lane, phase);
        float f[kNF];
        if (i < k.n) features_from(fk, cur, phase, f);
        else
#pragma unroll
            for (int q = 0; q < kNF; ++q) f[q] = 0.0f;
        float y0, y1;
        forward_regs(&L, f, b40, b41, y0, y1);
        if (i < k.n) {
            // The guess is kappa in units of the inverse Jacobi diagonal; for a
            // particle with almost no neighbours alpha (and so that unit) is up to
            // 13 orders larger than in the bulk, where the network's small output
            // error would become a huge pressure.  The unit is taken with alpha
            // capped at 10x a full neighbourhood's, so the guess stays bounded
            // there; the solve brings such a particle's pressure to its value.
            const float rho = cur.rho;
            const float cbase = fminf(cur.alpha, k.alpha_cap) * rho * rho * rho;
            const float c = phase ? cbase / (dt * dt) : cbase / dt;
            const float p0 = kappa_of(phase ? y1 : y0) * 1e-3f * c;
            const float p = isfinite(p0) ? p0 : 0.0f;
            b.p[i] = p;
            static_cast<float4*>(b.pos4)[i].w = p / fmaxf(rho * rho, 1.0e-5f);   // p / rho^2
        }
        cur = nxt;
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
    features_from(feat_k(k, b.status->dt, b.status->dt_prev), load_feat_in(b, i, phase), phase, f);
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
                                                  uint32_t nrows, float inv_batch, float tau1,
                                                  float* __restrict__ part, float* __restrict__ part_loss) {
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
    // Squared error over the two outputs, mean over the batch.  The density
    // output's error is weighted 2 tau when the prediction is low and
    // 2 (1 - tau) when high (an expectile loss; tau = 0.5 is plain MSE).
    float d0 = 0.0f, d1 = 0.0f;
    if (valid) {
        d0 = y[0] - m.ds_t[size_t(row) * kNO];
        d1 = y[1] - m.ds_t[size_t(row) * kNO + 1];
    }
    const float w1 = d1 < 0.0f ? 2.0f * tau1 : 2.0f * (1.0f - tau1);
    float l2 = d0 * d0 + w1 * d1 * d1;
    dy[2 * s] = 2.0f * d0 * inv_batch;
    dy[2 * s + 1] = 2.0f * w1 * d1 * inv_batch;
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
constexpr int kPackThreads = kRegChunks * 64 * 8 > kNH * LH ? kRegChunks * 64 * 8 : kNH * LH;

__global__ void __launch_bounds__(256) k_mlp_pack(MlpDev m) {
    WPackAll* all = static_cast<WPackAll*>(m.wpack);
    WPack* w = &all->lds;
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
    if (t < uint32_t(kRegChunks * 64 * 8))
        all->lreg[t] = bf16(lreg_value(m.theta, int(t >> 9), int((t >> 3) & 63u), int(t & 7u)));
}

// Forward on caller-supplied rows (evaluation and tests): out[r][0..1].  The
// same register-resident forward as the warm start.
__global__ void __launch_bounds__(64) k_mlp_eval(MlpDev m, const float* __restrict__ xin, uint32_t n,
                                                 float* __restrict__ out) {
    __shared__ InferLds L;
    load_infer(m, &L);
    __syncthreads();
    const uint32_t r = blockIdx.x * 64u + threadIdx.x;
    float f[kNF];
#pragma unroll
    for (int c = 0; c < kNF; ++c) f[c] = r < n ? xin[size_t(r) * kNF + c] : 0.0f;
    float y0, y1;
    forward_regs(&L, f, m.theta[kOB4], m.theta[kOB4 + 1], y0, y1);
    if (r < n) {
        out[2 * size_t(r)] = y0;
        out[2 * size_t(r) + 1] = y1;
    }
}

inline uint32_t blocks_for(uint64_t n, uint32_t t) { return uint32_t((n + t - 1) / t); }
inline hipStream_t S(void* s) { return static_cast<hipStream_t>(s); }

}  // namespace

size_t mlp_pack_bytes() { return sizeof(WPackAll); }

void launch_mlp_warm(const SimConst& k, const Buffers& b, const MlpDev& m, int phase, void* stream) {
    if (!k.n) return;
    // Three workgroups of four waves per CU: the kernel takes about 150 VGPRs
    // (three waves per SIMD) and 23 KB of LDS per workgroup.
    const uint32_t tiles = (k.n + 63u) / 64u;
    const uint32_t blocks = std::min(blocks_for(tiles, kInferThreads / 64), 3u * std::max(m.cus, 1u));
    DFSPH_LAUNCH(k_mlp_warm, dim3(blocks), dim3(kInferThreads), 0, S(stream), k, b, m, phase);
}
void launch_record_features(const SimConst& k, const Buffers& b, const MlpDev& m, int phase,
                            uint32_t step, uint32_t stride, uint32_t base, uint32_t cap, void* stream) {
    if (!k.n) return;
    DFSPH_LAUNCH(k_record_features, dim3(blocks_for(k.n, 256)), dim3(256), 0, S(stream), k, b, m,
                       phase, step, stride, base, cap);
}
void launch_record_targets(const SimConst& k, const Buffers& b, const MlpDev& m, uint32_t step,
                           uint32_t stride, uint32_t base, uint32_t cap, void* stream) {
    if (!k.n) return;
    DFSPH_LAUNCH(k_record_targets, dim3(blocks_for(k.n, 256)), dim3(256), 0, S(stream), k, b, m,
                       step, stride, base, cap);
}
void launch_mlp_pack(const MlpDev& m, void* stream) {
    DFSPH_LAUNCH(k_mlp_pack, dim3(blocks_for(kPackThreads, 256)), dim3(256), 0, S(stream), m);
}
void launch_mlp_train(const MlpDev& m, const uint32_t* rows, uint32_t nrows, float* part, float* part_loss,
                      float tau_density, void* stream) {
    DFSPH_LAUNCH(k_mlp_train, dim3(blocks_for(nrows, 64)), dim3(64), kTrainSmem, S(stream), m, rows,
                       nrows, 1.0f / float(nrows), tau_density, part, part_loss);
}
void launch_mlp_adam(const MlpDev& m, const float* part, uint32_t nparts, float lr, float b1, float b2,
                     float eps, float c1, float c2, float* grad_out, void* stream) {
    DFSPH_LAUNCH(k_mlp_adam, dim3(blocks_for(kNParam, 256)), dim3(256), 0, S(stream), m, part, nparts,
                       lr, b1, b2, eps, c1, c2, grad_out);
}
void launch_mlp_eval(const MlpDev& m, const float* x, uint32_t n, float* out, void* stream) {
    if (!n) return;
    DFSPH_LAUNCH(k_mlp_eval, dim3(blocks_for(n, 64)), dim3(64), 0, S(stream), m, x, n, out);
}
int mlp_set_smem_limits() {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(k_mlp_train),
                               hipFuncAttributeMaxDynamicSharedMemorySize, kTrainSmem) == hipSuccess
               ? 1
               : 0;
}
int mlp_smem_ok() { return kTrainSmem <= 160 * 1024 ? 1 : 0; }

}  // namespace dfsph
