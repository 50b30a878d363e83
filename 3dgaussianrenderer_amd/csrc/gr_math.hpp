// gr_math.hpp -- scalar/vector math for the splat pipeline, device and host.
//
// Every operation is written out in the order in which the reference's GLM
// 0.9.9.8 evaluates it, because tile lists must be reproduced bit for bit:
//   dot(vec2) = x*x' + y*y'                    (glm/detail/func_geometric.inl:37-44)
//   dot(vec3) = (x*x' + y*y') + z*z'           (func_geometric.inl:46-54)
//   normalize(v) = v * (1 / sqrt(dot(v, v)))   (func_geometric.inl:82-89, func_exponential.inl:136-139)
//   min(x, y) = (y < x) ? y : x, max(x, y) = (x < y) ? y : x     (func_common.inl:17-30)
//   clamp(x, lo, hi) = min(max(x, lo), hi)     (func_common.inl:240-246)
//   step(e, x) = (x < e) ? 0 : 1               (func_common.inl:249-255, 543-548)
//   mix(x, y, a) = x * (1 - a) + y * a         (func_common.inl:104-112)
//   mat3 * mat3: R[i][j] = (A[0][j]*B[i][0] + A[1][j]*B[i][1]) + A[2][j]*B[i][2]
//                                              (detail/type_mat3x3.inl:484-518)
//   mat4 * vec4: (m0*v0 + m1*v1) + (m2*v2 + m3*v3)   (detail/type_mat4x4.inl:536-570)
//   mat3_cast(quat)                            (gtc/quaternion.inl:41-66)
// All files that include this header are compiled with -ffp-contract=off, so
// no product is fused into an add unless an explicit fmaf() says so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GR_HD __host__ __device__ __forceinline__

namespace gr {

struct v2 { float x, y; };
struct v3 { float x, y, z; };
struct v4 { float x, y, z, w; };
// column-major like glm: c[i] is column i
struct m3 { v3 c[3]; };
struct m4 { v4 c[4]; };

GR_HD v2 mk2(float x, float y) { return v2{x, y}; }
GR_HD v3 mk3(float x, float y, float z) { return v3{x, y, z}; }

GR_HD float gmin(float x, float y) { return (y < x) ? y : x; }
GR_HD float gmax(float x, float y) { return (x < y) ? y : x; }
GR_HD float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
GR_HD int imin(int x, int y) { return (y < x) ? y : x; }
GR_HD int imax(int x, int y) { return (x < y) ? y : x; }
GR_HD int iclamp(int x, int lo, int hi) { return imin(imax(x, lo), hi); }
GR_HD float gstep(float edge, float x) { return (x < edge) ? 0.0f : 1.0f; }

// __saturatef of the reference (core/gaussian_render.cu:245): [0,1], NaN -> 0.
GR_HD float saturate(float x) { return (x > 0.0f) ? ((x < 1.0f) ? x : 1.0f) : 0.0f; }

// Saturating float -> int truncation, NaN -> 0 (what v_cvt_i32_f32 does; a
// bare C++ cast is undefined out of range, so both sides use this helper).
GR_HD int32_t f2i_sat(float x) {
    if (!(x == x)) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)x;
}
// Saturating float -> uint32 truncation, NaN -> 0.
GR_HD uint32_t f2u_sat(float x) {
    if (!(x == x)) return 0u;
    if (x >= 4294967296.0f) return 0xffffffffu;
    if (!(x > 0.0f)) return 0u;
    return (uint32_t)x;
}

GR_HD float dot2(v2 a, v2 b) { return a.x * b.x + a.y * b.y; }
GR_HD float dot3(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
GR_HD float len2sq(v2 a) { return dot2(a, a); }
GR_HD v2 normalize2(v2 v) {
    float s = 1.0f / sqrtf(dot2(v, v));
    return v2{v.x * s, v.y * s};
}
GR_HD v3 normalize3(v3 v) {
    float s = 1.0f / sqrtf(dot3(v, v));
    return v3{v.x * s, v.y * s, v.z * s};
}
GR_HD v3 cross3(v3 x, v3 y) {
    return v3{x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}

GR_HD float m3at(const m3& m, int c, int r) {
    const v3& col = m.c[c];
    return r == 0 ? col.x : (r == 1 ? col.y : col.z);
}
GR_HD m3 mul33(const m3& a, const m3& b) {
    m3 r;
    const float A00 = a.c[0].x, A01 = a.c[0].y, A02 = a.c[0].z;
    const float A10 = a.c[1].x, A11 = a.c[1].y, A12 = a.c[1].z;
    const float A20 = a.c[2].x, A21 = a.c[2].y, A22 = a.c[2].z;
    const float B00 = b.c[0].x, B01 = b.c[0].y, B02 = b.c[0].z;
    const float B10 = b.c[1].x, B11 = b.c[1].y, B12 = b.c[1].z;
    const float B20 = b.c[2].x, B21 = b.c[2].y, B22 = b.c[2].z;
    r.c[0].x = A00 * B00 + A10 * B01 + A20 * B02;
    r.c[0].y = A01 * B00 + A11 * B01 + A21 * B02;
    r.c[0].z = A02 * B00 + A12 * B01 + A22 * B02;
    r.c[1].x = A00 * B10 + A10 * B11 + A20 * B12;
    r.c[1].y = A01 * B10 + A11 * B11 + A21 * B12;
    r.c[1].z = A02 * B10 + A12 * B11 + A22 * B12;
    r.c[2].x = A00 * B20 + A10 * B21 + A20 * B22;
    r.c[2].y = A01 * B20 + A11 * B21 + A21 * B22;
    r.c[2].z = A02 * B20 + A12 * B21 + A22 * B22;
    return r;
}
GR_HD m3 transpose3(const m3& m) {
    m3 t;
    t.c[0] = v3{m.c[0].x, m.c[1].x, m.c[2].x};
    t.c[1] = v3{m.c[0].y, m.c[1].y, m.c[2].y};
    t.c[2] = v3{m.c[0].z, m.c[1].z, m.c[2].z};
    return t;
}
// mat4 * vec4, pairwise sum.
GR_HD v4 mul44v(const m4& m, v4 v) {
    v4 r;
    r.x = (m.c[0].x * v.x + m.c[1].x * v.y) + (m.c[2].x * v.z + m.c[3].x * v.w);
    r.y = (m.c[0].y * v.x + m.c[1].y * v.y) + (m.c[2].y * v.z + m.c[3].y * v.w);
    r.z = (m.c[0].z * v.x + m.c[1].z * v.y) + (m.c[2].z * v.z + m.c[3].z * v.w);
    r.w = (m.c[0].w * v.x + m.c[1].w * v.y) + (m.c[2].w * v.z + m.c[3].w * v.w);
    return r;
}
GR_HD m3 upper3(const m4& m) {
    m3 r;
    for (int i = 0; i < 3; ++i) r.c[i] = v3{m.c[i].x, m.c[i].y, m.c[i].z};
    return r;
}
GR_HD m4 load_m4(const float* p) {
    m4 m;
    for (int i = 0; i < 4; ++i) m.c[i] = v4{p[4 * i + 0], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]};
    return m;
}

// mat3_cast of glm::quat(w, x, y, z)
GR_HD m3 quat_to_m3(float w, float x, float y, float z) {
    const float qxx = x * x, qyy = y * y, qzz = z * z;
    const float qxz = x * z, qxy = x * y, qyz = y * z;
    const float qwx = w * x, qwy = w * y, qwz = w * z;
    m3 r;
    r.c[0].x = 1.0f - 2.0f * (qyy + qzz);
    r.c[0].y = 2.0f * (qxy + qwz);
    r.c[0].z = 2.0f * (qxz - qwy);
    r.c[1].x = 2.0f * (qxy - qwz);
    r.c[1].y = 1.0f - 2.0f * (qxx + qzz);
    r.c[1].z = 2.0f * (qyz + qwx);
    r.c[2].x = 2.0f * (qxz + qwy);
    r.c[2].y = 2.0f * (qyz - qwx);
    r.c[2].z = 1.0f - 2.0f * (qxx + qyy);
    return r;
}

// ---------------------------------------------------------------------------
// Frame geometry.  The reference fixes a 1024x1024 frame of 16x16 tiles
// (core/consts.h:4-8).  The per-axis generalisation reduces to it exactly at
// 1024x1024: every constant below is then the reference's constant.
// ---------------------------------------------------------------------------
constexpr int kTileSize = 16;

struct FrameGeom {
    int width, height;
    int tiles_x, tiles_y;        // ceil(W/16), ceil(H/16)
    float tiles_per_axis_x;      // W / 16.0f   (reference: (float)k_TilesPerScreen)
    float tiles_per_axis_y;      // H / 16.0f
    float tile_clip_x;           // (16 / (float)W) * 2   (core/gaussian_render.cu:471-472)
    float tile_clip_y;
    float texel_clip_x;          // 2 / (float)W          (core/gaussian_render.cu:185, 704)
    float texel_clip_y;
    float trace_bump_x;          // ((1/pi) * texel_x) * texel_x  (core/gaussian_render.cu:184-186)
    float trace_bump_y;
};

GR_HD FrameGeom make_geom(int width, int height) {
    FrameGeom g;
    g.width = width;
    g.height = height;
    g.tiles_x = (width + kTileSize - 1) / kTileSize;
    g.tiles_y = (height + kTileSize - 1) / kTileSize;
    g.tiles_per_axis_x = (float)width / (float)kTileSize;
    g.tiles_per_axis_y = (float)height / (float)kTileSize;
    g.tile_clip_x = ((float)kTileSize / (float)width) * 2.0f;
    g.tile_clip_y = ((float)kTileSize / (float)height) * 2.0f;
    g.texel_clip_x = 2.0f / (float)width;
    g.texel_clip_y = 2.0f / (float)height;
    const float kPi = 3.14159265359f;
    g.trace_bump_x = ((1.0f / kPi) * g.texel_clip_x) * g.texel_clip_x;
    g.trace_bump_y = ((1.0f / kPi) * g.texel_clip_y) * g.texel_clip_y;
    return g;
}

GR_HD int tile_bits_for(int tiles) {
    int b = 0;
    while ((1 << b) < tiles) ++b;
    return b;
}

// ---------------------------------------------------------------------------
// exp2 used by the blend.  The reference calls CUDA's __expf (closed code, its
// bits are not reproducible here).  This build defines density = 2^e with its
// own polynomial so that the GPU and the host oracle agree bit for bit:
//   n = floor(e), f = min(e - n, 1 - 2^-24) in [0, 1) (e - n rounded to nearest
//   even), 2^f ~ p(f), p(0) = 1 exactly, a degree-4 near-minimax relative-error
//   fit on [0, 1] (Lawson-weighted least squares on Chebyshev nodes,
//   tests/golden/fit_exp2.py; max relative error 3.0e-6 with FP32 Horner/FMA
//   evaluation), result = ldexp(p(f), n).  p(f) < 2 for every f < 1, so the
//   density never exceeds 1 for e <= 0, and 2^n is exact at integers.
// On gfx950 f is one v_fract_f32 (the same rounding and clamp) and n one
// v_cvt_flr_i32_f32.  Callers pass e in [-26, 0] (see blend_alpha_dev in the
// raster), so n fits an int without saturation.
// ---------------------------------------------------------------------------
constexpr float kExp2C0 = 1.0f;
constexpr float kExp2C1 = 6.930448413e-01f;
constexpr float kExp2C2 = 2.412801981e-01f;
constexpr float kExp2C3 = 5.224247277e-02f;
constexpr float kExp2C4 = 1.342668477e-02f;

__device__ __forceinline__ float exp2_poly(float e) {
    const float f = __builtin_amdgcn_fractf(e);
    int n;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(n) : "v"(e));
    float p = fmaf(kExp2C4, f, kExp2C3);
    p = fmaf(p, f, kExp2C2);
    p = fmaf(p, f, kExp2C1);
    p = fmaf(p, f, kExp2C0);
    return ldexpf(p, n);
}

// Blend weights.  conic (a, b, c) of the reference becomes
//   qa = a * (-0.5*log2 e), qb = b * (-log2 e), qc = c * (-0.5*log2 e)
// so that e = dx*(qa*dx + qb*dy) + (qc*dy)*dy = -0.5*log2(e)*power, with
// power = a dx^2 + c dy^2 + 2 b dx dy of core/gaussian_render.cu:732.
// alpha = opacity * 2^min(e, 0) when e >= -25, else 0 (a pair whose density
// is below 2^-25 contributes nothing; NaN contributes nothing).
constexpr float kHalfLog2e = -0.72134752044448170f;   // -0.5 * log2(e)
constexpr float kLog2e = -1.44269504088896340f;       // -log2(e)
constexpr float kMinExponent = -25.0f;
constexpr float kTransmittanceStop = 0.02f;           // core/gaussian_render.cu:745

GR_HD float blend_exponent(float qa, float qb, float qc, float dx, float dy) {
    return fmaf(dx, fmaf(qa, dx, qb * dy), (qc * dy) * dy);
}

}  // namespace gr
