// render.hip -- screen-space fluid render passes as compute kernels, replacing
// the graphics passes B-E of main.cpp:2900-3132:
//   project      particle_sprite.vert.glsl / thickness.vert.glsl: view and clip
//                position, window position, point sizes; points outside the
//                clip volume are dropped, as Vulkan clips a point primitive
//   depth_splat  particle_sprite.frag.glsl coverage: a pixel is covered when
//                its centre lies in the sprite's square and in its inscribed
//                disc (gl_PointCoord); the reference's depth attachment has no
//                depth test, so the fragment of the last draw -- the highest
//                particle index -- stays: atomicMax of the original index
//   depth_resolve that winner's sphere-impostor NDC z (1 where nothing drew)
//   bilateral    bilateral_smooth.comp.glsl, one direction per dispatch
//   thick_splat  thickness.frag.glsl Gaussian profile, additively blended:
//                fixed-point 64-bit sums (2^-32 units) so the result does not
//                depend on the order the atomics land in
//   composite    composite.frag.glsl: normal from fine screen-space
//                derivatives of the reconstructed position, Schlick Fresnel,
//                Beer-Lambert body colour, reflection/refraction of a two-
//                colour sky
// Built without FMA contraction: every expression rounds as written, in the
// order written, so a host restatement of the passes can follow it exactly.
#include "dfsph_internal.h"
#include "launch.h"

namespace dfsph {
namespace {

constexpr float kNoDepth = 0.99999f;    // bilateral_smooth.comp.glsl:26, composite.frag.glsl:43
constexpr double kThickUnit = 4294967296.0;   // fixed-point scale of the thickness sums
constexpr float kMaxPointSize = 8192.0f;      // point size limit of the device the reference targets

struct V3 { float x, y, z; };
struct V4 { float x, y, z, w; };

// Column-major m times (x, y, z, w), each row summed left to right.
__device__ __forceinline__ V4 mat_vec(const float* m, float x, float y, float z, float w) {
    V4 r;
    r.x = ((m[0] * x + m[4] * y) + m[8] * z) + m[12] * w;
    r.y = ((m[1] * x + m[5] * y) + m[9] * z) + m[13] * w;
    r.z = ((m[2] * x + m[6] * y) + m[10] * z) + m[14] * w;
    r.w = ((m[3] * x + m[7] * y) + m[11] * z) + m[15] * w;
    return r;
}
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 scale(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 normalize(V3 a) { return scale(a, 1.0f / sqrtf(dot(a, a))); }
__device__ __forceinline__ float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
__device__ __forceinline__ float mix(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// gl_PointCoord of the pixel with lower-left corner (px, py) for a point at
// window position (xw, yw) of size s, mapped to [-1, 1]; false when the pixel
// centre is outside the sprite's square.
__device__ __forceinline__ bool sprite_uv(int px, int py, float xw, float yw, float s, float& u, float& v) {
    const float sc = 0.5f + ((float(px) + 0.5f) - xw) / s;
    const float tc = 0.5f + ((float(py) + 0.5f) - yw) / s;
    if (!(sc >= 0.0f && sc < 1.0f && tc >= 0.0f && tc < 1.0f)) return false;
    u = sc * 2.0f - 1.0f;
    v = tc * 2.0f - 1.0f;
    return true;
}

// proj4[id] = (xw, yw, depth-pass size, thickness-pass size);
// view4[id] = (view-space x, y, z, 1 if the point survives clipping).
__global__ void __launch_bounds__(256) k_render_project(RenderConst k, const float4* __restrict__ x4,
                                                        const uint32_t* __restrict__ ids, float4* proj4,
                                                        float4* view4) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= k.n) return;
    const uint32_t id = ids[s];
    if (id >= k.n) return;
    const float4 p = x4[s];
    const V4 v = mat_vec(k.view, p.x, p.y, p.z, 1.0f);
    const V4 c = mat_vec(k.proj, v.x, v.y, v.z, v.w);
    const bool inside = c.w > 0.0f && -c.w <= c.x && c.x <= c.w && -c.w <= c.y && c.y <= c.w && 0.0f <= c.z &&
                        c.z <= c.w;
    const float half_w = float(k.w) * 0.5f, half_h = float(k.h) * 0.5f;
    const float xw = (c.x / c.w) * half_w + half_w;
    const float yw = (c.y / c.w) * half_h + half_h;
    const float vz = -v.z;
    const float size = 2.0f * k.radius * k.point_scale / fmaxf(vz, 0.01f);
    const float sd = fminf(fmaxf(size, 2.0f), 1023.0f);          // particle_sprite.vert.glsl:33
    const float st = fminf(fmaxf(2.0f, size), kMaxPointSize);     // thickness.vert.glsl:29, device limit
    proj4[id] = make_float4(xw, yw, sd, st);
    view4[id] = make_float4(v.x, v.y, v.z, inside ? 1.0f : 0.0f);
}

template <class F>
__device__ __forceinline__ void for_sprite_pixels(const RenderConst& k, float xw, float yw, float s, F&& fn) {
    const int x0 = max(0, int(floorf(xw - 0.5f * s)) - 1), x1 = min(int(k.w) - 1, int(floorf(xw + 0.5f * s)) + 1);
    const int y0 = max(0, int(floorf(yw - 0.5f * s)) - 1), y1 = min(int(k.h) - 1, int(floorf(yw + 0.5f * s)) + 1);
    for (int py = y0; py <= y1; ++py)
        for (int px = x0; px <= x1; ++px) {
            float u, v;
            if (!sprite_uv(px, py, xw, yw, s, u, v)) continue;
            const float r2 = u * u + v * v;
            if (r2 > 1.0f) continue;   // discard
            fn(px, py, r2);
        }
}

__global__ void __launch_bounds__(256) k_render_depth_splat(RenderConst k, const float4* __restrict__ proj4,
                                                            const float4* __restrict__ view4, int* winner) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= k.n || view4[id].w == 0.0f) return;
    const float4 q = proj4[id];
    for_sprite_pixels(k, q.x, q.y, q.z,
                      [&](int px, int py, float) { atomicMax(&winner[size_t(py) * k.w + px], int(id)); });
}

__global__ void __launch_bounds__(256) k_render_thick_splat(RenderConst k, const float4* __restrict__ proj4,
                                                            const float4* __restrict__ view4,
                                                            unsigned long long* acc) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= k.n) return;
    const float4 vp = view4[id];
    if (vp.w == 0.0f) return;
    const float4 q = proj4[id];
    const float inv_z = 1.0f / fmaxf(-vp.z, 0.1f);
    for_sprite_pixels(k, q.x, q.y, q.w, [&](int px, int py, float r2) {
        const float profile = expf(-r2 * 2.0f);
        const float t = k.tpp * profile * inv_z;
        atomicAdd(&acc[size_t(py) * k.w + px], __double2ull_rn(double(t) * kThickUnit));
    });
}

__global__ void __launch_bounds__(256) k_render_depth_resolve(RenderConst k, const float4* __restrict__ proj4,
                                                              const float4* __restrict__ view4,
                                                              const int* __restrict__ winner, float* depth,
                                                              const unsigned long long* __restrict__ acc,
                                                              float* thick) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= k.w * k.h) return;
    thick[p] = float(double(acc[p]) * (1.0 / kThickUnit));
    const int id = winner[p];
    if (id < 0) { depth[p] = 1.0f; return; }   // the attachment's clear value
    const int px = int(p % k.w), py = int(p / k.w);
    const float4 q = proj4[id];
    const float4 vp = view4[id];
    float u = 0.0f, v = 0.0f;
    sprite_uv(px, py, q.x, q.y, q.z, u, v);
    const float r2 = u * u + v * v;
    const float z_offset = sqrtf(1.0f - r2) * k.radius;
    const V4 c = mat_vec(k.proj, vp.x + u * k.radius, vp.y + v * k.radius, vp.z + z_offset, 1.0f);
    depth[p] = c.z / c.w;
}

__global__ void __launch_bounds__(256) k_render_bilateral(RenderConst k, const float* __restrict__ in, float* out,
                                                          int dir) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= k.w * k.h) return;
    const int px = int(p % k.w), py = int(p / k.w);
    const float center = in[p];
    if (center >= kNoDepth) { out[p] = center; return; }
    const int radius = int(ceilf(2.0f * k.sigma_s));
    const float inv_s = 1.0f / (2.0f * k.sigma_s * k.sigma_s);
    const float inv_d = 1.0f / (2.0f * k.sigma_d * k.sigma_d);
    float wsum = 0.0f, dsum = 0.0f;
    for (int dr = -radius; dr <= radius; ++dr) {
        const int sx = dir == 0 ? px + dr : px, sy = dir == 0 ? py : py + dr;
        if (sx < 0 || sy < 0 || sx >= int(k.w) || sy >= int(k.h)) continue;
        const float d = in[size_t(sy) * k.w + sx];
        if (d >= kNoDepth) continue;
        const float ws = expf(-float(dr * dr) * inv_s);
        const float dd = d - center;
        const float wd = expf(-(dd * dd) * inv_d);
        const float w = ws * wd;
        dsum += d * w;
        wsum += w;
    }
    out[p] = wsum > 1e-6f ? dsum / wsum : center;
}

__device__ __forceinline__ V3 reconstruct(const RenderConst& k, float u, float v, float ndc_z) {
    const V4 w = mat_vec(k.inv_vp, u * 2.0f - 1.0f, v * 2.0f - 1.0f, ndc_z, 1.0f);
    return {w.x / w.w, w.y / w.w, w.z / w.w};
}
__device__ __forceinline__ V3 sky(const RenderConst& k, V3 dir) {
    const float t = clamp01(dir.y * 0.5f + 0.5f);
    const float e = clamp01((t - 0.0f) / (0.7f - 0.0f));
    const float s = e * e * (3.0f - 2.0f * e);   // smoothstep(0, 0.7, t)
    return {mix(k.horizon[0], k.zenith[0], s), mix(k.horizon[1], k.zenith[1], s), mix(k.horizon[2], k.zenith[2], s)};
}
// The position a pixel's fragment reconstructs (helper lanes past the image
// edge sample the edge texel, as a clamped sampler does, at their own uv).
__device__ __forceinline__ V3 pixel_pos(const RenderConst& k, const float* depth, int px, int py) {
    const int cx = min(px, int(k.w) - 1), cy = min(py, int(k.h) - 1);
    const float u = (float(px) + 0.5f) / float(k.w), v = (float(py) + 0.5f) / float(k.h);
    return reconstruct(k, u, v, depth[size_t(cy) * k.w + cx]);
}

__global__ void __launch_bounds__(256) k_render_composite(RenderConst k, const float* __restrict__ depth,
                                                          const float* __restrict__ thick, float4* rgba) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= k.w * k.h) return;
    const int px = int(p % k.w), py = int(p / k.w);
    const float u = (float(px) + 0.5f) / float(k.w), v = (float(py) + 0.5f) / float(k.h);
    const V3 cam = {k.cam[0], k.cam[1], k.cam[2]};
    const float d = depth[p];
    if (d >= kNoDepth) {
        const V3 c = sky(k, normalize(sub(reconstruct(k, u, v, 1.0f), cam)));
        rgba[p] = make_float4(c.x * k.exposure, c.y * k.exposure, c.z * k.exposure, 1.0f);
        return;
    }
    // Fine derivatives: differences across the pixel's 2x2 quad.
    const V3 P = reconstruct(k, u, v, d);
    const V3 dpdx = sub(pixel_pos(k, depth, px | 1, py), pixel_pos(k, depth, px & ~1, py));
    const V3 dpdy = sub(pixel_pos(k, depth, px, py | 1), pixel_pos(k, depth, px, py & ~1));
    const V3 N = normalize(cross(dpdy, dpdx));
    const V3 V = normalize(sub(cam, P));
    const float cos_t = fmaxf(0.0f, dot(N, V));
    const float thickness = thick[p] * k.thick_scale;
    const float tint[3] = {k.tint[0], k.tint[1], k.tint[2]};
    float trans[3];
    for (int a = 0; a < 3; ++a) trans[a] = expf(-k.absorption * thickness * (1.0f - tint[a]));
    const V3 I = scale(V, -1.0f);
    const float ni = dot(N, I);
    const float eta = 1.0f / 1.33f;
    const float kk = 1.0f - eta * eta * (1.0f - ni * ni);
    V3 refr = {0.0f, 0.0f, 0.0f};
    if (!(kk < 0.0f)) refr = sub(scale(I, eta), scale(N, eta * ni + sqrtf(kk)));
    const V3 refl = sub(I, scale(N, 2.0f * ni));
    const V3 refr_c = sky(k, refr), refl_c = sky(k, refl);
    const float F = k.f0 + (1.0f - k.f0) * powf(clamp01(1.0f - cos_t), 5.0f);
    const float rc[3] = {refr_c.x, refr_c.y, refr_c.z}, lc[3] = {refl_c.x, refl_c.y, refl_c.z};
    const float zen[3] = {k.zenith[0], k.zenith[1], k.zenith[2]};
    float out[3];
    for (int a = 0; a < 3; ++a) {
        const float body = rc[a] * trans[a] + tint[a] * (1.0f - trans[a]);
        const float col = mix(body, lc[a], F);
        const float rough = mix(lc[a], zen[a], k.rough);
        out[a] = mix(col, rough, k.rough * 0.3f) * k.exposure;
    }
    rgba[p] = make_float4(out[0], out[1], out[2], 1.0f);
}

template <class T>
__global__ void __launch_bounds__(256) k_render_fill(T* dst, size_t count, T value) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = value;
}

inline unsigned blocks(size_t n) { return unsigned((n + 255) / 256); }

}  // namespace

const float* launch_render(const RenderConst& k, const RenderBuffers& b, int iterations, int alternate,
                           void* stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t pixels = size_t(k.w) * k.h;
    float4* proj4 = static_cast<float4*>(b.proj4);
    float4* view4 = static_cast<float4*>(b.view4);
    DFSPH_LAUNCH(k_render_fill<int>, dim3(blocks(pixels)), dim3(256), 0, s, b.winner, pixels, -1);
    DFSPH_LAUNCH(k_render_fill<unsigned long long>, dim3(blocks(pixels)), dim3(256), 0, s, b.thick_acc, pixels,
                 0ull);
    if (k.n) {
        DFSPH_LAUNCH(k_render_fill<float4>, dim3(blocks(k.n)), dim3(256), 0, s, view4, size_t(k.n),
                     make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        DFSPH_LAUNCH(k_render_project, dim3(blocks(k.n)), dim3(256), 0, s, k,
                     static_cast<const float4*>(b.x4), b.ids, proj4, view4);
        DFSPH_LAUNCH(k_render_depth_splat, dim3(blocks(k.n)), dim3(256), 0, s, k, proj4, view4, b.winner);
        DFSPH_LAUNCH(k_render_thick_splat, dim3(blocks(k.n)), dim3(256), 0, s, k, proj4, view4, b.thick_acc);
    }
    DFSPH_LAUNCH(k_render_depth_resolve, dim3(blocks(pixels)), dim3(256), 0, s, k, proj4, view4, b.winner,
                 b.depth, b.thick_acc, b.thick);
    // Iteration 0 reads the depth image, later ones ping-pong (main.cpp:2972-2996).
    const float* src = b.depth;
    for (int i = 0; i < iterations; ++i) {
        float* dst = (i % 2 == 0) ? b.smooth_a : b.smooth_b;
        DFSPH_LAUNCH(k_render_bilateral, dim3(blocks(pixels)), dim3(256), 0, s, k, src, dst,
                     alternate ? (i & 1) : 0);
        src = dst;
    }
    DFSPH_LAUNCH(k_render_composite, dim3(blocks(pixels)), dim3(256), 0, s, k, src, b.thick,
                 static_cast<float4*>(b.rgba));
    return src;
}

}  // namespace dfsph
