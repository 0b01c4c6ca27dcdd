// render_oracle.cpp -- CPU oracle for the screen-space fluid render passes.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/ as part of liboracle.so, never by
// the product.  It restates passes B-E of the reference's frame
// (particle-fluids/sph-water/src/main.cpp:2900-3132) the way the graphics
// pipeline runs them, in draw order:
//   * depth pass      shaders/particle_sprite.vert.glsl:23-33 and
//                     particle_sprite.frag.glsl:29-40 over every particle in
//                     index order; the colour attachment is cleared to 1 and has
//                     no depth test or blending (main.cpp:1692-1702), so each
//                     fragment overwrites the pixel: the last draw stays
//   * smoothing       shaders/bilateral_smooth.comp.glsl:21-57, iterations
//                     ping-ponged as main.cpp:2972-2996, every one along x
//                     (passDirection packed once as 0, main.cpp:2973) unless
//                     alternate is set
//   * thickness pass  shaders/thickness.vert.glsl:20-30 and
//                     thickness.frag.glsl:23-31, additive (ONE, ONE) blending in
//                     draw order (main.cpp:1704-1720)
//   * composite       shaders/composite.frag.glsl:29-78 over a full-screen
//                     triangle, derivatives taken as fine differences across
//                     each 2x2 quad; the thickness scale is
//                     absorption_thickness_pad.y (0 in main.cpp:2186)
// Point rasterisation: a point outside the clip volume is dropped; a pixel is
// covered when its centre lies in the square of side gl_PointSize around the
// window position (gl_PointCoord in [0, 1)), the fragment discarded when
// |2 gl_PointCoord - 1|^2 > 1.  Float arithmetic in the shaders' order,
// compiled without contraction.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct V3 { float x, y, z; };
struct V4 { float x, y, z, w; };

V4 mat_vec(const float* m, float x, float y, float z, float w) {
    return {((m[0] * x + m[4] * y) + m[8] * z) + m[12] * w, ((m[1] * x + m[5] * y) + m[9] * z) + m[13] * w,
            ((m[2] * x + m[6] * y) + m[10] * z) + m[14] * w, ((m[3] * x + m[7] * y) + m[11] * z) + m[15] * w};
}
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 scale(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V3 normalize(V3 a) { return scale(a, 1.0f / std::sqrt(dot(a, a))); }
float clamp01(float x) { return std::min(std::max(x, 0.0f), 1.0f); }
float mix(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// glm::inverse-equivalent cofactor inverse of a column-major 4x4 (expansion by
// complementary 2x2 minors of the top and bottom row pairs).
void inverse4(const float* m, float* out) {
    auto A = [&](int r, int c) { return m[c * 4 + r]; };
    const float s0 = A(0, 0) * A(1, 1) - A(1, 0) * A(0, 1), s1 = A(0, 0) * A(1, 2) - A(1, 0) * A(0, 2);
    const float s2 = A(0, 0) * A(1, 3) - A(1, 0) * A(0, 3), s3 = A(0, 1) * A(1, 2) - A(1, 1) * A(0, 2);
    const float s4 = A(0, 1) * A(1, 3) - A(1, 1) * A(0, 3), s5 = A(0, 2) * A(1, 3) - A(1, 2) * A(0, 3);
    const float c5 = A(2, 2) * A(3, 3) - A(3, 2) * A(2, 3), c4 = A(2, 1) * A(3, 3) - A(3, 1) * A(2, 3);
    const float c3 = A(2, 1) * A(3, 2) - A(3, 1) * A(2, 2), c2 = A(2, 0) * A(3, 3) - A(3, 0) * A(2, 3);
    const float c1 = A(2, 0) * A(3, 2) - A(3, 0) * A(2, 2), c0 = A(2, 0) * A(3, 1) - A(3, 0) * A(2, 1);
    const float det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
    const float id = 1.0f / det;
    auto I = [&](int r, int c, float v) { out[c * 4 + r] = v * id; };
    I(0, 0, A(1, 1) * c5 - A(1, 2) * c4 + A(1, 3) * c3);
    I(0, 1, -A(0, 1) * c5 + A(0, 2) * c4 - A(0, 3) * c3);
    I(0, 2, A(3, 1) * s5 - A(3, 2) * s4 + A(3, 3) * s3);
    I(0, 3, -A(2, 1) * s5 + A(2, 2) * s4 - A(2, 3) * s3);
    I(1, 0, -A(1, 0) * c5 + A(1, 2) * c2 - A(1, 3) * c1);
    I(1, 1, A(0, 0) * c5 - A(0, 2) * c2 + A(0, 3) * c1);
    I(1, 2, -A(3, 0) * s5 + A(3, 2) * s2 - A(3, 3) * s1);
    I(1, 3, A(2, 0) * s5 - A(2, 2) * s2 + A(2, 3) * s1);
    I(2, 0, A(1, 0) * c4 - A(1, 1) * c2 + A(1, 3) * c0);
    I(2, 1, -A(0, 0) * c4 + A(0, 1) * c2 - A(0, 3) * c0);
    I(2, 2, A(3, 0) * s4 - A(3, 1) * s2 + A(3, 3) * s0);
    I(2, 3, -A(2, 0) * s4 + A(2, 1) * s2 - A(2, 3) * s0);
    I(3, 0, -A(1, 0) * c3 + A(1, 1) * c1 - A(1, 2) * c0);
    I(3, 1, A(0, 0) * c3 - A(0, 1) * c1 + A(0, 2) * c0);
    I(3, 2, -A(3, 0) * s3 + A(3, 1) * s1 - A(3, 2) * s0);
    I(3, 3, A(2, 0) * s3 - A(2, 1) * s1 + A(2, 2) * s0);
}

struct Frame {
    const float *view, *proj;
    uint32_t w, h;
    float radius, point_scale, tpp;
};

// One point primitive: clip test, window position, gl_PointSize (both
// passes' clamps), view position.
struct Point {
    bool drawn;
    float xw, yw, size_depth, size_thick;
    V3 view;
};
Point vertex(const Frame& f, const float* x) {
    const V4 v = mat_vec(f.view, x[0], x[1], x[2], 1.0f);
    const V4 c = mat_vec(f.proj, v.x, v.y, v.z, v.w);
    Point p{};
    p.drawn = c.w > 0.0f && -c.w <= c.x && c.x <= c.w && -c.w <= c.y && c.y <= c.w && 0.0f <= c.z && c.z <= c.w;
    const float hw = float(f.w) * 0.5f, hh = float(f.h) * 0.5f;
    p.xw = (c.x / c.w) * hw + hw;
    p.yw = (c.y / c.w) * hh + hh;
    const float vz = -v.z;
    const float size = 2.0f * f.radius * f.point_scale / std::max(vz, 0.01f);
    p.size_depth = std::min(std::max(size, 2.0f), 1023.0f);
    p.size_thick = std::min(std::max(2.0f, size), 8192.0f);
    p.view = {v.x, v.y, v.z};
    return p;
}

// Fragments of a point of size s: fn(pixel, u, v, r2) for every covered,
// undiscarded pixel.
template <class F>
void rasterize(const Frame& f, float xw, float yw, float s, F&& fn) {
    const int x0 = std::max(0, int(std::floor(xw - 0.5f * s)) - 2);
    const int x1 = std::min(int(f.w) - 1, int(std::floor(xw + 0.5f * s)) + 2);
    const int y0 = std::max(0, int(std::floor(yw - 0.5f * s)) - 2);
    const int y1 = std::min(int(f.h) - 1, int(std::floor(yw + 0.5f * s)) + 2);
    for (int py = y0; py <= y1; ++py)
        for (int px = x0; px <= x1; ++px) {
            const float sc = 0.5f + ((float(px) + 0.5f) - xw) / s;   // gl_PointCoord
            const float tc = 0.5f + ((float(py) + 0.5f) - yw) / s;
            if (!(sc >= 0.0f && sc < 1.0f && tc >= 0.0f && tc < 1.0f)) continue;
            const float u = sc * 2.0f - 1.0f, v = tc * 2.0f - 1.0f;
            const float r2 = u * u + v * v;
            if (r2 > 1.0f) continue;
            fn(size_t(py) * f.w + size_t(px), u, v, r2);
        }
}

V3 sky(const float* horizon, const float* zenith, V3 dir) {
    const float t = clamp01(dir.y * 0.5f + 0.5f);
    const float e = clamp01((t - 0.0f) / (0.7f - 0.0f));
    const float s = e * e * (3.0f - 2.0f * e);
    return {mix(horizon[0], zenith[0], s), mix(horizon[1], zenith[1], s), mix(horizon[2], zenith[2], s)};
}

}  // namespace

extern "C" {

// Bilateral smoothing (bilateral_smooth.comp.glsl:21-57), `iterations`
// dispatches ping-ponged (main.cpp:2972-2996), along x unless alternate.
void oracle_render_bilateral(const float* depth_in, uint32_t w, uint32_t h, int iterations, float sigma_s,
                             float sigma_d, int alternate, float* out) {
    const size_t px = size_t(w) * h;
    std::vector<float> depth(depth_in, depth_in + px);
    std::vector<float> a(px), b(px);
    const std::vector<float>* src = &depth;
    for (int it = 0; it < iterations; ++it) {
        std::vector<float>& dst = (it % 2 == 0) ? a : b;
        const int dir = alternate ? (it & 1) : 0;
        const std::vector<float>& in = *src;
        const int rad = int(std::ceil(2.0f * sigma_s));
        const float inv_s = 1.0f / (2.0f * sigma_s * sigma_s);
        const float inv_d = 1.0f / (2.0f * sigma_d * sigma_d);
        for (uint32_t py = 0; py < h; ++py)
            for (uint32_t pxl = 0; pxl < w; ++pxl) {
                const size_t q = size_t(py) * w + pxl;
                const float center = in[q];
                if (center >= 0.99999f) { dst[q] = center; continue; }
                float ws_sum = 0.0f, d_sum = 0.0f;
                for (int dr = -rad; dr <= rad; ++dr) {
                    const int sx = dir == 0 ? int(pxl) + dr : int(pxl), sy = dir == 0 ? int(py) : int(py) + dr;
                    if (sx < 0 || sy < 0 || sx >= int(w) || sy >= int(h)) continue;
                    const float d = in[size_t(sy) * w + size_t(sx)];
                    if (d >= 0.99999f) continue;
                    const float wsp = std::exp(-float(dr * dr) * inv_s);
                    const float dd = d - center;
                    const float wdp = std::exp(-(dd * dd) * inv_d);
                    const float wgt = wsp * wdp;
                    d_sum += d * wgt;
                    ws_sum += wgt;
                }
                dst[q] = ws_sum > 1e-6f ? d_sum / ws_sum : center;
            }
        src = &dst;
    }
    std::memcpy(out, src->data(), px * 4);
}

// Composite (composite.frag.glsl:29-78) of a smoothed depth and a thickness
// image into rgba (x4).
void oracle_render_composite(const float* fin, const float* thick, uint32_t w, uint32_t h, const float* view,
                             const float* proj, const float* cam, const float* tint, float rough,
                             const float* zenith, const float* horizon, float exposure, float f0,
                             float absorption, float thick_scale, float* rgba) {
    float vp[16], ivp[16];
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r)
            vp[c * 4 + r] = ((proj[r] * view[c * 4] + proj[4 + r] * view[c * 4 + 1]) + proj[8 + r] * view[c * 4 + 2]) +
                            proj[12 + r] * view[c * 4 + 3];
    inverse4(vp, ivp);
    const V3 C = {cam[0], cam[1], cam[2]};
    auto recon = [&](float u, float v, float z) {
        const V4 q = mat_vec(ivp, u * 2.0f - 1.0f, v * 2.0f - 1.0f, z, 1.0f);
        return V3{q.x / q.w, q.y / q.w, q.z / q.w};
    };
    auto pos_at = [&](int qx, int qy) {   // a quad member's reconstruction (clamped texel, own uv)
        const int cx = std::min(qx, int(w) - 1), cy = std::min(qy, int(h) - 1);
        return recon((float(qx) + 0.5f) / float(w), (float(qy) + 0.5f) / float(h), fin[size_t(cy) * w + size_t(cx)]);
    };
    for (uint32_t py = 0; py < h; ++py)
        for (uint32_t pxl = 0; pxl < w; ++pxl) {
            const size_t q = size_t(py) * w + pxl;
            const float u = (float(pxl) + 0.5f) / float(w), v = (float(py) + 0.5f) / float(h);
            float out[3];
            const float d = fin[q];
            if (d >= 0.99999f) {
                const V3 s = sky(horizon, zenith, normalize(sub(recon(u, v, 1.0f), C)));
                out[0] = s.x * exposure; out[1] = s.y * exposure; out[2] = s.z * exposure;
            } else {
                const int ix = int(pxl), iy = int(py);
                const V3 P = recon(u, v, d);
                const V3 dpdx = sub(pos_at(ix | 1, iy), pos_at(ix & ~1, iy));
                const V3 dpdy = sub(pos_at(ix, iy | 1), pos_at(ix, iy & ~1));
                const V3 N = normalize(cross(dpdy, dpdx));
                const V3 V = normalize(sub(C, P));
                const float cos_t = std::max(0.0f, dot(N, V));
                const float thickness = thick[q] * thick_scale;
                float trans[3];
                for (int k = 0; k < 3; ++k) trans[k] = std::exp(-absorption * thickness * (1.0f - tint[k]));
                const V3 I = scale(V, -1.0f);
                const float ni = dot(N, I);
                const float eta = 1.0f / 1.33f;
                const float kk = 1.0f - eta * eta * (1.0f - ni * ni);
                V3 refr = {0.0f, 0.0f, 0.0f};
                if (!(kk < 0.0f)) refr = sub(scale(I, eta), scale(N, eta * ni + std::sqrt(kk)));
                const V3 refl = sub(I, scale(N, 2.0f * ni));
                const V3 rc = sky(horizon, zenith, refr), lc = sky(horizon, zenith, refl);
                const float F = f0 + (1.0f - f0) * std::pow(clamp01(1.0f - cos_t), 5.0f);
                const float rcs[3] = {rc.x, rc.y, rc.z}, lcs[3] = {lc.x, lc.y, lc.z};
                for (int k = 0; k < 3; ++k) {
                    const float body = rcs[k] * trans[k] + tint[k] * (1.0f - trans[k]);
                    const float col = mix(body, lcs[k], F);
                    const float rough_sky = mix(lcs[k], zenith[k], rough);
                    out[k] = mix(col, rough_sky, rough * 0.3f) * exposure;
                }
            }
            if (rgba) {
                rgba[4 * q] = out[0]; rgba[4 * q + 1] = out[1]; rgba[4 * q + 2] = out[2]; rgba[4 * q + 3] = 1.0f;
            }
        }
}

// Positions n x 3 in draw (original index) order.  Outputs width*height
// (rgba x4), row 0 at the top.
void oracle_render(const float* x, uint32_t n, const float* view, const float* proj, const float* cam,
                   uint32_t w, uint32_t h, float radius, float point_scale, float tpp, int iterations,
                   float sigma_s, float sigma_d, int alternate, const float* tint, float rough,
                   const float* zenith, const float* horizon, float exposure, float f0, float absorption,
                   float thick_scale, float* rgba, float* depth_out, float* smoothed_out, float* thick_out) {
    const Frame f{view, proj, w, h, radius, point_scale, tpp};
    const size_t px = size_t(w) * h;
    std::vector<Point> pts(n);
    for (uint32_t i = 0; i < n; ++i) pts[i] = vertex(f, x + 3 * size_t(i));

    // Depth pass: each fragment overwrites the pixel (draw order).
    std::vector<float> depth(px, 1.0f);
    for (uint32_t i = 0; i < n; ++i) {
        const Point& p = pts[i];
        if (!p.drawn) continue;
        rasterize(f, p.xw, p.yw, p.size_depth, [&](size_t q, float u, float v, float r2) {
            const float z_offset = std::sqrt(1.0f - r2) * radius;
            const V4 c = mat_vec(proj, p.view.x + u * radius, p.view.y + v * radius, p.view.z + z_offset, 1.0f);
            depth[q] = c.z / c.w;
        });
    }
    // Thickness pass: additive blending in draw order.
    std::vector<float> thick(px, 0.0f);
    for (uint32_t i = 0; i < n; ++i) {
        const Point& p = pts[i];
        if (!p.drawn) continue;
        const float inv_z = 1.0f / std::max(-p.view.z, 0.1f);
        rasterize(f, p.xw, p.yw, p.size_thick, [&](size_t q, float, float, float r2) {
            thick[q] = thick[q] + tpp * std::exp(-r2 * 2.0f) * inv_z;
        });
    }
    std::vector<float> fin(px);
    oracle_render_bilateral(depth.data(), w, h, iterations, sigma_s, sigma_d, alternate, fin.data());
    oracle_render_composite(fin.data(), thick.data(), w, h, view, proj, cam, tint, rough, zenith, horizon,
                            exposure, f0, absorption, thick_scale, rgba);
    if (depth_out) std::memcpy(depth_out, depth.data(), px * 4);
    if (smoothed_out) std::memcpy(smoothed_out, fin.data(), px * 4);
    if (thick_out) std::memcpy(thick_out, thick.data(), px * 4);
}

}  // extern "C"
