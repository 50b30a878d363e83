// oracle/mirror.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference's per-frame splat pipeline
// (core/gaussian_render.cu), against which the HIP kernels are compared.  Only
// the tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// it; the product never links, imports or calls anything here.
//
// Two renders live here:
//   * the FP32 mirror (orc_*), which follows the reference kernels operation
//     by operation in IEEE binary32 without contraction (built with
//     -ffp-contract=off, no fast-math), so that its SH colours, projections,
//     tile lists and sort order can be compared bit for bit with the GPU;
//   * the FP64 render (orc64_*), the same algorithm in binary64 with the
//     reference's own formulas (atan2 -> cos/sin, exp, T *= 1 - alpha), used
//     as the accuracy yardstick (PSNR >= 50 dB).
//
// Where the reference cannot be reproduced bit for bit (it uses CUDA's
// __expf and nvcc's default FMA contraction, closed code whose results are not
// specified), the FP32 mirror follows the arithmetic this build defines:
//   - (cos, sin) of the ellipse = the normalised eigenvector itself instead
//     of cos/sin(atan2(e0)) (core/gaussian_render.cu:200,212);
//   - blend: e = dx*(qa*dx + qb*dy) + (qc*dy)*dy with qa = a*(-0.5 log2 e),
//     qb = b*(-log2 e), qc = c*(-0.5 log2 e) evaluated with two FMAs; density
//     2^clamp(e,-26,0) from a degree-5 polynomial + ldexp; alpha = 0 when
//     e < -25 (or NaN);
//     w = T*alpha, C = fma(rgb, w, C), T = T - w.
// Parity pins for this file live in tests/test_oracle.py (closed forms,
// brute force, FP64 agreement, textbook sort) -- see the header of that file.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <type_traits>
#include <vector>

#include <omp.h>

#define ORC_API extern "C" __attribute__((visibility("default")))

namespace {

// ---- GLM semantics (external/glm/glm/detail/func_common.inl:17-30,240-255) ----
template <typename T> inline T gmin(T x, T y) { return (y < x) ? y : x; }
template <typename T> inline T gmax(T x, T y) { return (x < y) ? y : x; }
template <typename T> inline T gclamp(T x, T lo, T hi) { return gmin(gmax(x, lo), hi); }
template <typename T> inline T gstep(T edge, T x) { return (x < edge) ? T(0) : T(1); }
inline int iclamp(int x, int lo, int hi) { return std::min(std::max(x, lo), hi); }
// __saturatef: [0,1], NaN -> 0
template <typename T> inline T sat(T x) { return (x > T(0)) ? ((x < T(1)) ? x : T(1)) : T(0); }

// saturating conversions (CUDA cvt.rzi semantics; NaN -> 0)
inline int32_t to_i32(double x) {
    if (!(x == x)) return 0;
    if (x >= 2147483648.0) return INT32_MAX;
    if (x <= -2147483648.0) return INT32_MIN;
    return (int32_t)x;
}
inline uint32_t to_u32(double x) {
    if (!(x == x)) return 0u;
    if (x >= 4294967296.0) return 0xffffffffu;
    if (!(x > 0.0)) return 0u;
    return (uint32_t)x;
}

// column-major 3x3 / 4x4, GLM product order
template <typename T> struct M3 { T m[3][3]; };  // m[col][row]
template <typename T> struct M4 { T m[4][4]; };

template <typename T> M3<T> mul(const M3<T>& a, const M3<T>& b) {
    // detail/type_mat3x3.inl:484-518
    M3<T> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.m[i][j] = a.m[0][j] * b.m[i][0] + a.m[1][j] * b.m[i][1] + a.m[2][j] * b.m[i][2];
    return r;
}
template <typename T> M3<T> tr(const M3<T>& a) {
    M3<T> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
    return r;
}
template <typename T> void mulv(const M4<T>& m, const T v[4], T out[4]) {
    // detail/type_mat4x4.inl:536-570: (m0*v0 + m1*v1) + (m2*v2 + m3*v3)
    for (int r = 0; r < 4; ++r)
        out[r] = (m.m[0][r] * v[0] + m.m[1][r] * v[1]) + (m.m[2][r] * v[2] + m.m[3][r] * v[3]);
}
template <typename T> M3<T> upper(const M4<T>& a) {
    M3<T> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j];
    return r;
}

// ---- frame geometry (1024x1024 in the reference, core/consts.h:4-8) ----
template <typename T> struct Geom {
    int W, H, TX, TY;
    T tpa_x, tpa_y, tclip_x, tclip_y, texel_x, texel_y, bump_x, bump_y;
};
template <typename T> Geom<T> geom(int W, int H) {
    Geom<T> g;
    g.W = W;
    g.H = H;
    g.TX = (W + 15) / 16;
    g.TY = (H + 15) / 16;
    g.tpa_x = (T)W / (T)16;
    g.tpa_y = (T)H / (T)16;
    g.tclip_x = ((T)16 / (T)W) * (T)2;   // core/gaussian_render.cu:471-472
    g.tclip_y = ((T)16 / (T)H) * (T)2;
    g.texel_x = (T)2 / (T)W;             // core/gaussian_render.cu:185, 704
    g.texel_y = (T)2 / (T)H;
    // core/gaussian_render.cu:184-186; the FP32 literal pi of the reference
    const T kPi = std::is_same<T, float>::value ? (T)3.14159265359f : (T)3.14159265358979323846;
    g.bump_x = (((T)1 / kPi) * g.texel_x) * g.texel_x;
    g.bump_y = (((T)1 / kPi) * g.texel_y) * g.texel_y;
    return g;
}

int tile_bits(int tiles) {
    int b = 0;
    while ((1 << b) < tiles) ++b;
    return b;
}

// ---- SH: core/gaussian_render.cu:50-116, band 3 core/cuda_renderer.cu:119-157 ----
template <typename T>
void sh_one(int degree, T x, T y, T z, const float* sh, int stride, T out[3]) {
    auto c = [&](int k) { return (T)sh[(size_t)k * stride]; };
    T res[3];
    for (int ch = 0; ch < 3; ++ch) res[ch] = T(0) + (T)0.282094792f * c(ch);
    if (degree > 0) {
        const T k1 = (T)0.488602512f * y, k2 = (T)0.488602512f * z, k3 = (T)0.488602512f * x;
        for (int ch = 0; ch < 3; ++ch) res[ch] = res[ch] + ((k1 * c(3 + ch) + k2 * c(6 + ch)) + k3 * c(9 + ch));
        if (degree > 1) {
            const T xx = x * x, yy = y * y;
            const T k4 = ((T)1.09254843f * x) * y;
            const T k5 = ((T)1.09254843f * y) * z;
            const T k6 = (((T)-0.946174696f * xx) - (T)0.946174696f * yy) + (T)0.630783131f;
            const T k7 = ((T)1.09254843f * x) * z;
            const T k8 = ((T)0.546274215f * (x - y)) * (x + y);
            for (int ch = 0; ch < 3; ++ch)
                res[ch] = res[ch] + ((((k4 * c(12 + ch) + k5 * c(15 + ch)) + k6 * c(18 + ch)) + k7 * c(21 + ch)) +
                                     k8 * c(24 + ch));
            if (degree > 2) {
                const T k9 = ((T)0.295021795f * y) * ((T)6 * xx - (T)2 * yy);
                const T k10 = (((T)2.89061144f * x) * y) * z;
                const T k11 = ((T)3.6563664f * y) * (((T)-0.625f * xx - (T)0.625f * yy) + (T)0.5f);
                const T k12 = ((T)0.373176333f * z) * (((T)-5 * xx - (T)5 * yy) + (T)2);
                const T k13 = ((T)0.457045799f * x) * (((T)-5 * xx - (T)5 * yy) + (T)4);
                const T k14 = (((T)1.44530572f * z) * (x - y)) * (x + y);
                const T k15 = ((T)0.59004359f * x) * (xx - (T)3 * yy);
                for (int ch = 0; ch < 3; ++ch)
                    res[ch] = res[ch] + ((((((k9 * c(27 + ch) + k10 * c(30 + ch)) + k11 * c(33 + ch)) +
                                            k12 * c(36 + ch)) + k13 * c(39 + ch)) + k14 * c(42 + ch)) +
                                         k15 * c(45 + ch));
            }
        }
    }
    for (int ch = 0; ch < 3; ++ch) out[ch] = gclamp(res[ch] + (T)0.5f, T(0), T(1));
}

// Reads the aligned layout [group][coef][256] (app/main.cpp:147-177,
// core/gaussian_render.cu:109-112).
template <typename T>
void sh_all(const float* pos, const float* sh_aligned, int n, int degree, const float* campos, T* color) {
    const int sh_count = 3 * (degree + 1) * (degree + 1);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const T vx = (T)campos[0] - (T)pos[4 * i + 0];
        const T vy = (T)campos[1] - (T)pos[4 * i + 1];
        const T vz = (T)campos[2] - (T)pos[4 * i + 2];
        const T s = T(1) / std::sqrt((vx * vx + vy * vy) + vz * vz);  // glm::normalize
        const T x = vx * s, y = vy * s, z = vz * s;
        const size_t grp = (size_t)i / 256, lane = (size_t)i % 256;
        const float* base = sh_aligned + grp * 256 * sh_count + lane;
        T rgb[3];
        sh_one<T>(degree, x, y, z, base, 256, rgb);
        color[4 * i + 0] = rgb[0];
        color[4 * i + 1] = rgb[1];
        color[4 * i + 2] = rgb[2];
        color[4 * i + 3] = (T)pos[4 * i + 3];
    }
}

// ---- projection: core/gaussian_render.cu:135-215 ----
struct Cam {
    float view[16];
    float fc[2];
    float dsb[2];
};

template <typename T>
void project_one(const float* p4, const float* sr4, const Cam& cam, const Geom<T>& g, T xy[2], T* zo,
                 T ell[4], T con[4]) {
    M3<T> S = {};
    S.m[0][0] = (T)sr4[0];
    S.m[1][1] = (T)sr4[1];
    S.m[2][2] = (T)sr4[2];
    uint32_t q8;
    std::memcpy(&q8, &sr4[3], 4);
    // decodeVec4 (core/gaussian_render.cu:45-48) * 2 - 1; glm::quat(w, x, y, z)
    const T rx = ((T)((q8 >> 24) & 255u) / (T)255) * (T)2 - (T)1;
    const T ry = ((T)((q8 >> 16) & 255u) / (T)255) * (T)2 - (T)1;
    const T rz = ((T)((q8 >> 8) & 255u) / (T)255) * (T)2 - (T)1;
    const T rw = ((T)(q8 & 255u) / (T)255) * (T)2 - (T)1;
    // mat3_cast (gtc/quaternion.inl:41-66)
    M3<T> R;
    {
        const T x = rx, y = ry, z = rz, w = rw;
        const T qxx = x * x, qyy = y * y, qzz = z * z, qxz = x * z, qxy = x * y, qyz = y * z;
        const T qwx = w * x, qwy = w * y, qwz = w * z;
        R.m[0][0] = T(1) - T(2) * (qyy + qzz);
        R.m[0][1] = T(2) * (qxy + qwz);
        R.m[0][2] = T(2) * (qxz - qwy);
        R.m[1][0] = T(2) * (qxy - qwz);
        R.m[1][1] = T(1) - T(2) * (qxx + qzz);
        R.m[1][2] = T(2) * (qyz + qwx);
        R.m[2][0] = T(2) * (qxz + qwy);
        R.m[2][1] = T(2) * (qyz - qwx);
        R.m[2][2] = T(1) - T(2) * (qxx + qyy);
    }
    const M3<T> RS = mul(R, S);
    const M3<T> sigma = mul(RS, tr(RS));

    M4<T> view;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) view.m[c][r] = (T)cam.view[4 * c + r];
    const T pin[4] = {(T)p4[0], (T)p4[1], (T)p4[2], T(1)};
    T v[4];
    mulv(view, pin, v);
    const T fcx = (T)cam.fc[0], fcy = (T)cam.fc[1];
    const T zr = T(1) / v[2];
    const T zr2 = zr * zr;
    const T sx = -fcx * zr, sy = -fcy * zr;
    const T ftx = fcx * v[0], fty = fcy * v[1];
    const T shx = ftx * zr2, shy = fty * zr2;
    const T tx = -ftx * zr, ty = -fty * zr;
    M4<T> A = {};
    for (int i = 0; i < 4; ++i) A.m[i][i] = T(1);
    A.m[0][0] = sx;
    A.m[1][1] = sy;
    A.m[2][2] = (T)cam.dsb[0];
    A.m[2][0] = shx;
    A.m[2][1] = shy;
    A.m[3][0] = tx;
    A.m[3][1] = ty;
    A.m[3][2] = (T)cam.dsb[1];
    const M3<T> VP = mul(upper(A), upper(view));
    M3<T> cc = mul(mul(VP, tr(sigma)), tr(VP));
    const T vin[4] = {v[0], v[1], v[2], T(1)};
    T clip4[4];
    mulv(A, vin, clip4);
    T clip[3] = {clip4[0], clip4[1], clip4[2]};
    cc.m[0][0] = cc.m[0][0] + g.bump_x;
    cc.m[1][1] = cc.m[1][1] + g.bump_y;
    const T a = cc.m[0][0], b = cc.m[1][0], c = cc.m[1][1];
    const T det = a * c - b * b;
    const T mid = (T)0.5f * (a + c);
    const T eps = (T)1e-12f;
    const T rad = std::sqrt(gmax(eps, mid * mid - det));
    const T l0 = mid + rad;
    const T l1 = gmax(T(0), mid - rad);
    // normalize(vec2(b, l0 - a))
    const T e0x = b, e0y = l0 - a;
    const T inv = T(1) / std::sqrt(e0x * e0x + e0y * e0y);
    T cs, sn;
    if (std::is_same<T, float>::value) {
        cs = e0x * inv;
        sn = e0y * inv;
    } else {
        const T ang = std::atan2(e0y * inv, e0x * inv);  // core/gaussian_render.cu:200,212
        cs = std::cos(ang);
        sn = std::sin(ang);
    }
    T ex = std::sqrt(l0) * (T)3, ey = std::sqrt(l1) * (T)3;
    const T invDet = T(1) / gmax(eps, det);
    const T cx = c * invDet, cy = (-b) * invDet, cz = a * invDet;
    T edge[3];
    for (int k = 0; k < 3; ++k) edge[k] = gstep(T(-1), clip[k]) * gstep(clip[k], T(1));
    const T vis = ((edge[0] * edge[1]) * edge[2]) * gstep(T(0), l1);
    for (int k = 0; k < 3; ++k) clip[k] = (T)-128 * (T(1) - vis) + clip[k] * vis;
    ex = ex * vis;
    ey = ey * vis;
    xy[0] = clip[0];
    xy[1] = clip[1];
    *zo = clip[2];
    ell[0] = cs;
    ell[1] = sn;
    ell[2] = ex;
    ell[3] = ey;
    con[0] = cx;
    con[1] = cy;
    con[2] = cz;
    con[3] = T(0);
}

// ---- footprint + exact test: core/gaussian_render.cu:234-298, 364-395, 460-490 ----
struct Rect {
    int x0, y0, nx, ny;
};

template <typename T>
Rect footprint(const T c[2], const T e[4], const Geom<T>& g) {
    const T rx = e[0] * e[2], ry = e[1] * e[2];
    const T ux = e[1] * e[3], uy = (-e[0]) * e[3];
    T mn[2] = {(T)1e12f, (T)1e12f}, mx[2] = {(T)-1e12f, (T)-1e12f};
    for (int j = 0; j != 4; ++j) {
        const int bottom = j & 1, top = j >> 1;
        const T r = (T)(bottom ^ top) * (T)2 - (T)1;
        const T u = (T)top * (T)2 - (T)1;
        const T vx = rx * r + ux * u, vy = ry * r + uy * u;
        mn[0] = gmin(mn[0], vx);
        mn[1] = gmin(mn[1], vy);
        mx[0] = gmax(mx[0], vx);
        mx[1] = gmax(mx[1], vy);
    }
    mn[0] = mn[0] + c[0];
    mn[1] = mn[1] + c[1];
    mx[0] = mx[0] + c[0];
    mx[1] = mx[1] + c[1];
    const T f0 = ((mn[0] + T(1)) * (T)0.5f) * g.tpa_x;
    const T f1 = ((mn[1] + T(1)) * (T)0.5f) * g.tpa_y;
    const T f2 = ((mx[0] + T(1)) * (T)0.5f) * g.tpa_x;
    const T f3 = ((mx[1] + T(1)) * (T)0.5f) * g.tpa_y;
    Rect r;
    r.x0 = iclamp(to_i32((double)std::floor(f0)), 0, g.TX);
    r.y0 = iclamp(to_i32((double)std::floor(f1)), 0, g.TY);
    const int x1 = iclamp(to_i32((double)std::ceil(f2)), 0, g.TX);
    const int y1 = iclamp(to_i32((double)std::ceil(f3)), 0, g.TY);
    r.nx = x1 - r.x0;
    r.ny = y1 - r.y0;
    return r;
}

template <typename T>
inline void to_ell(T px, T py, const T c[2], const T e[4], T out[2]) {
    const T dx = px - c[0], dy = py - c[1];
    out[0] = (dx * e[0] + dy * e[1]) / e[2];
    out[1] = (dy * e[0] - dx * e[1]) / e[3];
}

template <typename T>
inline bool unit_circle(const T v0[2], const T v1[2]) {
    const T dx = v1[0] - v0[0], dy = v1[1] - v0[1];
    const T lsq = dx * dx + dy * dy;
    const T t = sat(((-v0[0]) * dx + (-v0[1]) * dy) / lsq);
    const T px = v0[0] + t * (v1[0] - v0[0]);
    const T py = v0[1] + t * (v1[1] - v0[1]);
    return (px * px + py * py) < T(1);
}

template <typename T>
bool overlap(int gx, int gy, const T c[2], const T e[4], const Geom<T>& g) {
    const T mnx = (T)gx * g.tclip_x - T(1), mny = (T)gy * g.tclip_y - T(1);
    const T mxx = mnx + g.tclip_x, mxy = mny + g.tclip_y;
    bool ov = c[0] > mnx && c[0] < mxx && c[1] > mny && c[1] < mxy;
    const T ctx = mnx + (mxx - mnx) * (T)0.5f, cty = mny + (mxy - mny) * (T)0.5f;
    T q[2];
    to_ell(ctx, cty, c, e, q);
    ov |= (q[0] * q[0] + q[1] * q[1]) < T(1);
    T p[4][2];
    to_ell(mnx, mny, c, e, p[0]);
    to_ell(mxx, mny, c, e, p[1]);
    to_ell(mxx, mxy, c, e, p[2]);
    to_ell(mnx, mxy, c, e, p[3]);
    ov |= unit_circle(p[0], p[1]);
    ov |= unit_circle(p[1], p[2]);
    ov |= unit_circle(p[2], p[3]);
    ov |= unit_circle(p[3], p[0]);
    return ov;
}

template <typename T>
inline uint32_t depth_bits(T z) {
    // getKey (core/gaussian_render.cu:306-310); uint32 max promotes to 2^32
    const T q = gclamp((z + T(1)) * (T)0.5f, T(0), T(1)) * (T)4294967296.0;
    return to_u32((double)q);
}

// Candidate tiles of one Gaussian in row-major order (core/gaussian_render.cu:466-470).
// A rect with nx <= 0 or ny <= 0 (only from NaN corners) yields nothing.
template <typename T, typename F>
void for_each_tile(const T c[2], const T e[4], const Geom<T>& g, F&& f) {
    const Rect r = footprint(c, e, g);
    if (r.nx <= 0 || r.ny <= 0) return;
    const int cnt = r.nx * r.ny;
    for (int t = 0; t < cnt; ++t) {
        const int gx = r.x0 + t % r.nx, gy = r.y0 + t / r.nx;
        if (overlap(gx, gy, c, e, g)) f(gx, gy);
    }
}

// Deterministic build: per-Gaussian counts, prefix sum, emission in Gaussian
// index order.  Writes min(K, capacity) entries; returns K.
template <typename T>
int64_t build(const T* xy, const T* z, const T* ell, int n, const Geom<T>& g, uint64_t* keys,
              int32_t* values, int64_t capacity) {
    std::vector<int64_t> cnt((size_t)n + 1, 0);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int i = 0; i < n; ++i) {
        int64_t k = 0;
        for_each_tile<T>(&xy[2 * i], &ell[4 * i], g, [&](int, int) { ++k; });
        cnt[i] = k;
    }
    int64_t run = 0;
    for (int i = 0; i < n; ++i) {
        const int64_t c = cnt[i];
        cnt[i] = run;
        run += c;
    }
    cnt[n] = run;
    if (capacity > 0 && keys && values) {
#pragma omp parallel for schedule(dynamic, 1024)
        for (int i = 0; i < n; ++i) {
            int64_t pos = cnt[i];
            const uint32_t d = depth_bits(z[i]);
            for_each_tile<T>(&xy[2 * i], &ell[4 * i], g, [&](int gx, int gy) {
                if (pos < capacity) {
                    keys[pos] = ((uint64_t)(uint32_t)(gy * g.TX + gx) << 32) | d;
                    values[pos] = i;
                }
                ++pos;
            });
        }
    }
    return run;
}

// Stable LSD radix sort on bits [begin, end), 8-bit digits, counting sort per
// digit (the semantics of cub::DeviceRadixSort::SortPairs used at
// core/gaussian_render.cu:600-621).
void radix_sort(uint64_t* keys, int32_t* vals, int64_t n, int begin, int end) {
    if (n <= 1) return;
    std::vector<uint64_t> k2((size_t)n);
    std::vector<int32_t> v2((size_t)n);
    uint64_t* ka = keys;
    int32_t* va = vals;
    uint64_t* kb = k2.data();
    int32_t* vb = v2.data();
    const int nt = omp_get_max_threads();
    std::vector<int64_t> hist((size_t)nt * 256);
    for (int shift = begin; shift < end; shift += 8) {
        const int bits = std::min(8, end - shift);
        const uint64_t mask = (1ull << bits) - 1ull;
        std::fill(hist.begin(), hist.end(), 0);
#pragma omp parallel num_threads(nt)
        {
            const int t = omp_get_thread_num();
            const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
            int64_t* h = &hist[(size_t)t * 256];
            for (int64_t i = lo; i < hi; ++i) ++h[(ka[i] >> shift) & mask];
        }
        int64_t run = 0;
        for (int d = 0; d < 256; ++d)
            for (int t = 0; t < nt; ++t) {
                const int64_t c = hist[(size_t)t * 256 + d];
                hist[(size_t)t * 256 + d] = run;
                run += c;
            }
#pragma omp parallel num_threads(nt)
        {
            const int t = omp_get_thread_num();
            const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
            int64_t* h = &hist[(size_t)t * 256];
            for (int64_t i = lo; i < hi; ++i) {
                const int64_t p = h[(ka[i] >> shift) & mask]++;
                kb[p] = ka[i];
                vb[p] = va[i];
            }
        }
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        std::memcpy(keys, ka, (size_t)n * 8);
        std::memcpy(vals, va, (size_t)n * 4);
    }
}

// core/gaussian_render.cu:634-660 with the -1 pre-fill of app/main.cpp:365
void ranges(const uint64_t* keys, int64_t k, int tiles, int32_t* range) {
    std::fill(range, range + 2 * (size_t)tiles, -1);
    if (k <= 0) return;
    range[2 * (keys[0] >> 32)] = 0;
    range[2 * (keys[k - 1] >> 32) + 1] = (int32_t)k;
    for (int64_t i = 1; i < k; ++i) {
        const uint32_t p = (uint32_t)(keys[i - 1] >> 32), c = (uint32_t)(keys[i] >> 32);
        if (p != c) {
            range[2 * (size_t)p + 1] = (int32_t)i;
            range[2 * (size_t)c] = (int32_t)i;
        }
    }
}

// ---- blend arithmetic of this build (FP32) ----
// 2^e as gr::exp2_poly (gr_math.hpp): n = floor(e), f = min(e - n, 1 - 2^-24) (what gfx950's
// v_fract_f32 returns), degree-4 polynomial of tests/golden/fit_exp2.py, ldexp.
constexpr float kE2[5] = {1.0f, 6.930448413e-01f, 2.412801981e-01f, 5.224247277e-02f, 1.342668477e-02f};
inline float exp2_poly(float e) {
    const float fl = std::floor(e);
    const float f = std::fmin(e - fl, 0x1.fffffep-1f);
    float p = std::fma(kE2[4], f, kE2[3]);
    p = std::fma(p, f, kE2[2]);
    p = std::fma(p, f, kE2[1]);
    p = std::fma(p, f, kE2[0]);
    return std::ldexp(p, (int)fl);
}

// Raster FP32 (core/gaussian_render.cu:679-759 with the blend above).
void raster32(const int32_t* range, const int32_t* vals, const float* xy, const float* con,
              const float* col, const Geom<float>& g, uint8_t* out) {
    const int tiles = g.TX * g.TY;
    const float kHalf = -0.72134752044448170f, kFull = -1.44269504088896340f;
#pragma omp parallel for schedule(dynamic, 4)
    for (int tile = 0; tile < tiles; ++tile) {
        const int tx = tile % g.TX, ty = tile / g.TX;
        const int start = range[2 * tile], end = range[2 * tile + 1];
        const int n = end - start;
        float C[256][3], T[256], pcx[256], pcy[256];
        for (int p = 0; p < 256; ++p) {
            C[p][0] = C[p][1] = C[p][2] = 0.0f;
            T[p] = 1.0f;
            pcx[p] = (float)(tx * 16 + p % 16) * g.texel_x - 1.0f;
            pcy[p] = (float)(ty * 16 + p / 16) * g.texel_y - 1.0f;
        }
        for (int b0 = 0; b0 < n; b0 += 32) {
            const int b1 = std::min(n, b0 + 32);
            for (int j = b0; j < b1; ++j) {
                const int gi = vals[start + j];
                const float cx = xy[2 * gi], cy = xy[2 * gi + 1];
                const float qa = con[4 * gi] * kHalf, qb = con[4 * gi + 1] * kFull, qc = con[4 * gi + 2] * kHalf;
                const float op = col[4 * gi + 3], r = col[4 * gi], gg = col[4 * gi + 1], bb = col[4 * gi + 2];
                for (int p = 0; p < 256; ++p) {
                    const float dx = pcx[p] - cx, dy = pcy[p] - cy;
                    const float e = std::fma(dx, std::fma(qa, dx, qb * dy), (qc * dy) * dy);
                    float ec = (e < 0.0f) ? e : 0.0f;
                    ec = (ec > -26.0f) ? ec : -26.0f;
                    float al = op * exp2_poly(ec);
                    al = (e >= -25.0f) ? al : 0.0f;
                    const float w = T[p] * al;
                    C[p][0] = std::fma(r, w, C[p][0]);
                    C[p][1] = std::fma(gg, w, C[p][1]);
                    C[p][2] = std::fma(bb, w, C[p][2]);
                    T[p] = T[p] - w;
                }
            }
            bool alive = false;
            for (int p = 0; p < 256; ++p) alive |= T[p] > 0.02f;
            if (!alive) break;
        }
        for (int p = 0; p < 256; ++p) {
            const int px = tx * 16 + p % 16, py = ty * 16 + p / 16;
            if (px >= g.W || py >= g.H) continue;
            uint8_t* o = out + 4 * ((size_t)py * g.W + px);
            if (n > 0) {
                for (int k = 0; k < 3; ++k) o[k] = (uint8_t)to_u32((double)(gclamp(C[p][k], 0.0f, 1.0f) * 255.0f));
                o[3] = 255;
            } else {
                o[0] = o[1] = o[2] = o[3] = 0;
            }
        }
    }
}

// Raster FP64 with the reference's own formula (core/gaussian_render.cu:727-737).
void raster64(const int32_t* range, const int32_t* vals, const double* xy, const double* con,
              const double* col, const Geom<double>& g, uint8_t* out, double* out_float) {
    const int tiles = g.TX * g.TY;
#pragma omp parallel for schedule(dynamic, 4)
    for (int tile = 0; tile < tiles; ++tile) {
        const int tx = tile % g.TX, ty = tile / g.TX;
        const int start = range[2 * tile], end = range[2 * tile + 1];
        const int n = end - start;
        double C[256][3], T[256], pcx[256], pcy[256];
        for (int p = 0; p < 256; ++p) {
            C[p][0] = C[p][1] = C[p][2] = 0.0;
            T[p] = 1.0;
            pcx[p] = (double)(tx * 16 + p % 16) * g.texel_x - 1.0;
            pcy[p] = (double)(ty * 16 + p / 16) * g.texel_y - 1.0;
        }
        for (int b0 = 0; b0 < n; b0 += 32) {
            const int b1 = std::min(n, b0 + 32);
            for (int j = b0; j < b1; ++j) {
                const int gi = vals[start + j];
                for (int p = 0; p < 256; ++p) {
                    const double dx = pcx[p] - xy[2 * gi], dy = pcy[p] - xy[2 * gi + 1];
                    const double pw = con[4 * gi] * dx * dx + con[4 * gi + 2] * dy * dy + 2.0 * con[4 * gi + 1] * dx * dy;
                    const double al = col[4 * gi + 3] * sat(std::exp(-0.5 * pw));
                    for (int k = 0; k < 3; ++k) C[p][k] += col[4 * gi + k] * T[p] * al;
                    T[p] *= (1.0 - al);
                }
            }
            bool alive = false;
            for (int p = 0; p < 256; ++p) alive |= T[p] > 0.02;
            if (!alive) break;
        }
        for (int p = 0; p < 256; ++p) {
            const int px = tx * 16 + p % 16, py = ty * 16 + p / 16;
            if (px >= g.W || py >= g.H) continue;
            const size_t o = (size_t)py * g.W + px;
            for (int k = 0; k < 3; ++k) {
                const double v = n > 0 ? gclamp(C[p][k], 0.0, 1.0) : 0.0;
                out[4 * o + k] = (uint8_t)to_u32(v * 255.0);
                if (out_float) out_float[3 * o + k] = v;
            }
            out[4 * o + 3] = n > 0 ? 255 : 0;
        }
    }
}

Cam make_cam(const float* cam) {
    Cam c;
    std::memcpy(c.view, cam, 16 * 4);
    c.fc[0] = cam[16];
    c.fc[1] = cam[17];
    c.dsb[0] = cam[18];
    c.dsb[1] = cam[19];
    return c;
}

}  // namespace

// ===========================================================================
// C-ABI of the oracle.  `cam` is packed as view[16] (column-major), fov
// cotangent[2], depth scale/bias[2], camera position[3] (23 floats).
// ===========================================================================
ORC_API void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
ORC_API int orc_get_threads(void) { return omp_get_max_threads(); }

// realignSphericalHarmonics (app/main.cpp:147-177); padding lanes zero.
ORC_API void orc_realign_sh(const float* src, float* dst, int group, int sh_count, int n) {
    const int groups = (n + group - 1) / group;
    for (int grp = 0; grp < groups; ++grp) {
        const size_t start = (size_t)grp * group * sh_count;
        const int here = std::min(group, n - grp * group);
        for (int i = 0; i < sh_count; ++i)
            for (int k = 0; k < group; ++k)
                dst[start + (size_t)group * i + k] = k < here ? src[start + (size_t)sh_count * k + i] : 0.0f;
    }
}

// encodeVec4 (core/ply_parser.cpp:231-239)
ORC_API uint32_t orc_encode_vec4(float x, float y, float z, float w) {
    const float v[4] = {x, y, z, w};
    uint32_t r = 0;
    for (int i = 0; i < 4; ++i) r |= (to_u32((double)(gclamp(v[i], 0.0f, 1.0f) * 255.0f)) & 255u) << (24 - 8 * i);
    return r;
}

ORC_API void orc_sh(const float* position, const float* sh_aligned, int n, int degree, const float* cam,
                    float* color) {
    sh_all<float>(position, sh_aligned, n, degree, cam + 20, color);
}

ORC_API void orc_project(const float* position, const float* scale_rot, int n, const float* cam, int W, int H,
                         float* xy, float* z, float* ell, float* con) {
    const Cam c = make_cam(cam);
    const Geom<float> g = geom<float>(W, H);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        project_one<float>(&position[4 * i], &scale_rot[4 * i], c, g, &xy[2 * i], &z[i], &ell[4 * i], &con[4 * i]);
}

ORC_API int64_t orc_build(const float* xy, const float* z, const float* ell, int n, int W, int H, uint64_t* keys,
                          int32_t* values, int64_t capacity) {
    return build<float>(xy, z, ell, n, geom<float>(W, H), keys, values, capacity);
}

// Tile counts per Gaussian (candidate rect and exact), for tests.
ORC_API void orc_tile_counts(const float* xy, const float* ell, int n, int W, int H, int32_t* rect_count,
                             int32_t* exact_count) {
    const Geom<float> g = geom<float>(W, H);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const Rect r = footprint<float>(&xy[2 * i], &ell[4 * i], g);
        rect_count[i] = (r.nx > 0 && r.ny > 0) ? r.nx * r.ny : 0;
        int k = 0;
        for_each_tile<float>(&xy[2 * i], &ell[4 * i], g, [&](int, int) { ++k; });
        exact_count[i] = k;
    }
}

ORC_API int orc_sort_end_bit(int W, int H) {
    const int tiles = ((W + 15) / 16) * ((H + 15) / 16);
    return 32 + tile_bits(tiles);
}

ORC_API void orc_sort(uint64_t* keys, int32_t* values, int64_t k, int begin_bit, int end_bit) {
    radix_sort(keys, values, k, begin_bit, end_bit);
}

ORC_API void orc_ranges(const uint64_t* keys, int64_t k, int tiles, int32_t* range) { ranges(keys, k, tiles, range); }

ORC_API void orc_raster(const int32_t* range, const int32_t* values, const float* xy, const float* con,
                        const float* col, int W, int H, uint8_t* out) {
    raster32(range, values, xy, con, col, geom<float>(W, H), out);
}

ORC_API float orc_exp2(float e) { return exp2_poly(e); }

// 16-bit storage of the precision form -> binary32 (both conversions exact).
static float half_to_float(uint16_t h) {
    const int sign = (h >> 15) & 1, e = (h >> 10) & 31, m = h & 1023;
    float v;
    if (e == 0) v = std::ldexp((float)m, -24);
    else if (e == 31) v = m ? std::numeric_limits<float>::quiet_NaN() : std::numeric_limits<float>::infinity();
    else v = std::ldexp((float)(m | 1024), e - 25);
    return sign ? -v : v;
}
static float bf16_to_float(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float v;
    std::memcpy(&v, &u, 4);
    return v;
}

// Decodes the precision form (position.w = quaternion bits, packed = scale
// x, y, z + opacity in 16 bits, SH aligned in 16 bits) into the FP32 form that
// the other oracle functions take; precision 1 = FP16, 2 = BF16.
ORC_API void orc_decode_precision(const float* position_p, const uint16_t* packed, const uint16_t* sh16, int n,
                                  int degree, int precision, float* position, float* scale_rot, float* sh_aligned) {
    auto dec = precision == 1 ? half_to_float : bf16_to_float;
    for (int i = 0; i < n; ++i) {
        position[4 * i + 0] = position_p[4 * i + 0];
        position[4 * i + 1] = position_p[4 * i + 1];
        position[4 * i + 2] = position_p[4 * i + 2];
        position[4 * i + 3] = dec(packed[4 * i + 3]);
        scale_rot[4 * i + 0] = dec(packed[4 * i + 0]);
        scale_rot[4 * i + 1] = dec(packed[4 * i + 1]);
        scale_rot[4 * i + 2] = dec(packed[4 * i + 2]);
        scale_rot[4 * i + 3] = position_p[4 * i + 3];
    }
    if (degree > 0) {
        const size_t total = (size_t)((n + 255) / 256) * 256 * 3 * (degree + 1) * (degree + 1);
        for (size_t k = 0; k < total; ++k) sh_aligned[k] = dec(sh16[k]);
    }
}

// Per-stage wall-clock (ms) of the last orc_render call.
static double g_stage_ms[6];
ORC_API void orc_last_stage_ms(double* out) { std::memcpy(out, g_stage_ms, sizeof(g_stage_ms)); }

// Whole FP32 frame (app/main.cpp:326-433 order).  sh_aligned may be NULL when
// degree == 0, in which case `color` must hold the upload colours.
// Scratch: caller passes buffers sized for n Gaussians; keys/values are
// allocated here.  Returns K; writes the image and, if non-NULL, the sorted
// tile list (first min(K, list_capacity) entries).
ORC_API int64_t orc_render(const float* position, const float* scale_rot, const float* sh_aligned, int n,
                           int degree, const float* cam, int W, int H, float* color, float* xy, float* z,
                           float* ell, float* con, uint8_t* image, uint64_t* list_keys, int32_t* list_values,
                           int64_t list_capacity) {
    auto now = [] { return omp_get_wtime() * 1000.0; };
    const double t0 = now();
    if (degree > 0) orc_sh(position, sh_aligned, n, degree, cam, color);
    const double t1 = now();
    orc_project(position, scale_rot, n, cam, W, H, xy, z, ell, con);
    const double t2 = now();
    const Geom<float> g = geom<float>(W, H);
    const int64_t k = build<float>(xy, z, ell, n, g, nullptr, nullptr, 0);
    std::vector<uint64_t> keys((size_t)std::max<int64_t>(k, 1));
    std::vector<int32_t> vals((size_t)std::max<int64_t>(k, 1));
    build<float>(xy, z, ell, n, g, keys.data(), vals.data(), k);
    const double t3 = now();
    radix_sort(keys.data(), vals.data(), k, 0, orc_sort_end_bit(W, H));
    const double t4 = now();
    std::vector<int32_t> range((size_t)2 * g.TX * g.TY);
    ranges(keys.data(), k, g.TX * g.TY, range.data());
    raster32(range.data(), vals.data(), xy, con, color, g, image);
    const double t5 = now();
    g_stage_ms[0] = t1 - t0;
    g_stage_ms[1] = t2 - t1;
    g_stage_ms[2] = t3 - t2;
    g_stage_ms[3] = t4 - t3;
    g_stage_ms[4] = t5 - t4;
    g_stage_ms[5] = t5 - t0;
    if (list_keys && list_values) {
        const int64_t m = std::min(k, list_capacity);
        std::memcpy(list_keys, keys.data(), (size_t)m * 8);
        std::memcpy(list_values, vals.data(), (size_t)m * 4);
    }
    return k;
}

// Whole FP64 frame with the reference's formulas (see file header).  Inputs
// are the FP32 scene (the reference's storage); everything after the load is
// binary64.  Returns K of the FP64 tile list; image_float (optional) gets the
// unquantised linear colour, 3 doubles per pixel.
ORC_API int64_t orc64_render(const float* position, const float* scale_rot, const float* sh_aligned,
                             const float* color_in, int n, int degree, const float* cam, int W, int H,
                             uint8_t* image, double* image_float, uint64_t* list_keys, int32_t* list_values,
                             int64_t list_capacity) {
    const Cam c = make_cam(cam);
    const Geom<double> g = geom<double>(W, H);
    std::vector<double> col((size_t)4 * n), xy((size_t)2 * n), z((size_t)n), ell((size_t)4 * n), con((size_t)4 * n);
    if (degree > 0) {
        sh_all<double>(position, sh_aligned, n, degree, cam + 20, col.data());
    } else {
        for (size_t i = 0; i < (size_t)4 * n; ++i) col[i] = color_in[i];
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        project_one<double>(&position[4 * i], &scale_rot[4 * i], c, g, &xy[2 * i], &z[i], &ell[4 * i], &con[4 * i]);
    const int64_t k = build<double>(xy.data(), z.data(), ell.data(), n, g, nullptr, nullptr, 0);
    std::vector<uint64_t> keys((size_t)std::max<int64_t>(k, 1));
    std::vector<int32_t> vals((size_t)std::max<int64_t>(k, 1));
    build<double>(xy.data(), z.data(), ell.data(), n, g, keys.data(), vals.data(), k);
    // stable sort by key == sort by (key, emission order)
    std::vector<int64_t> idx((size_t)k);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return keys[a] < keys[b]; });
    std::vector<uint64_t> sk((size_t)std::max<int64_t>(k, 1));
    std::vector<int32_t> sv((size_t)std::max<int64_t>(k, 1));
    for (int64_t i = 0; i < k; ++i) {
        sk[i] = keys[idx[i]];
        sv[i] = vals[idx[i]];
    }
    std::vector<int32_t> range((size_t)2 * g.TX * g.TY);
    ranges(sk.data(), k, g.TX * g.TY, range.data());
    raster64(range.data(), sv.data(), xy.data(), con.data(), col.data(), g, image, image_float);
    if (list_keys && list_values) {
        const int64_t m = std::min(k, list_capacity);
        std::memcpy(list_keys, sk.data(), (size_t)m * 8);
        std::memcpy(list_values, sv.data(), (size_t)m * 4);
    }
    return k;
}

// FP64 SH / projection of single inputs, for closed-form tests.
ORC_API void orc64_project(const float* position, const float* scale_rot, int n, const float* cam, int W, int H,
                           double* xy, double* z, double* ell, double* con) {
    const Cam c = make_cam(cam);
    const Geom<double> g = geom<double>(W, H);
    for (int i = 0; i < n; ++i)
        project_one<double>(&position[4 * i], &scale_rot[4 * i], c, g, &xy[2 * i], &z[i], &ell[4 * i], &con[4 * i]);
}

ORC_API void orc64_sh(const float* position, const float* sh_aligned, int n, int degree, const float* cam,
                      double* color) {
    sh_all<double>(position, sh_aligned, n, degree, cam + 20, color);
}
