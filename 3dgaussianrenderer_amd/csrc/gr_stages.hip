// gr_stages.hip -- per-frame stage kernels for gfx950 (MI355X, wave64).
//
//   SH colour      one lane per Gaussian, coefficients read from the aligned
//                  [group][coef][256] layout so that each wave load is 256
//                  contiguous bytes (HBM-bound).
//   projection     one lane per Gaussian, float4 SoA loads (HBM-bound).
//   build          deterministic count -> scan -> emit; one wave walks the
//                  candidate tiles of 64 Gaussians, 64 candidates at a time,
//                  and emits by ballot/popcount so no atomics decide order.
//   ranges         per-entry boundary test on sorted keys.
//   raster         one 256-lane workgroup per 16x16 tile (four 8x8 wave
//                  quadrants), entries staged through LDS 256 at a time,
//                  front-to-back blending, tile-wide stop vote every 32
//                  entries (the reference's rule).
#include <algorithm>

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "gr_kernels.hpp"
#include "gr_overlap.hpp"

namespace gr {

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t lanemask_lt(int lane) { return (1ull << lane) - 1ull; }

// ---------------------------------------------------------------------------
// SH colour: core/gaussian_render.cu:50-116 (bands 0-2), band 3 from
// core/cuda_renderer.cu:119-157.  Sums are taken band by band, each band's
// terms left to right, then added to the running colour, as the reference
// source is written.
// ---------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ float sh_ld(const T* p);
template <> __device__ __forceinline__ float sh_ld<float>(const float* p) {
    return __builtin_nontemporal_load(p);
}
template <> __device__ __forceinline__ float sh_ld<__half>(const __half* p) {
    return __half2float(*p);
}
template <> __device__ __forceinline__ float sh_ld<uint16_t>(const uint16_t* p) {
    // bf16 stored as its upper 16 bits
    return __uint_as_float(((uint32_t)*p) << 16);
}

// 16-bit halves of the precision form -> float (exact)
template <typename T> __device__ __forceinline__ float h16(uint32_t bits);
template <> __device__ __forceinline__ float h16<__half>(uint32_t bits) {
    return __half2float(__ushort_as_half((unsigned short)bits));
}
template <> __device__ __forceinline__ float h16<uint16_t>(uint32_t bits) {
    return __uint_as_float(bits << 16);
}

// Colour of one Gaussian seen from camera position cp: the reference's band sums
// (core/gaussian_render.cu:62-92, band 3 core/cuda_renderer.cu:119-157), clamped, with the
// given opacity in w.
template <int DEG>
__device__ __forceinline__ float4 sh_color(const float (&c)[3 * (DEG + 1) * (DEG + 1)], const float4& p, v3 cp,
                                           float opacity) {
    const v3 dir = normalize3(v3{cp.x - p.x, cp.y - p.y, cp.z - p.z});
    const float x = dir.x, y = dir.y, z = dir.z;
    float res[3];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) res[ch] = 0.0f + 0.282094792f * c[ch];
    if (DEG > 0) {
        const float k1 = 0.488602512f * y, k2 = 0.488602512f * z, k3 = 0.488602512f * x;
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            const float t = (k1 * c[3 + ch] + k2 * c[6 + ch]) + k3 * c[9 + ch];
            res[ch] = res[ch] + t;
        }
        if (DEG > 1) {
            const float xx = x * x, yy = y * y;
            const float k4 = (1.09254843f * x) * y;
            const float k5 = (1.09254843f * y) * z;
            const float k6 = ((-0.946174696f * xx) - 0.946174696f * yy) + 0.630783131f;
            const float k7 = (1.09254843f * x) * z;
            const float k8 = (0.546274215f * (x - y)) * (x + y);
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
                const float t = (((k4 * c[12 + ch] + k5 * c[15 + ch]) + k6 * c[18 + ch]) + k7 * c[21 + ch]) +
                                k8 * c[24 + ch];
                res[ch] = res[ch] + t;
            }
            if (DEG > 2) {
                const float k9 = (0.295021795f * y) * (6.0f * xx - 2.0f * yy);
                const float k10 = ((2.89061144f * x) * y) * z;
                const float k11 = (3.6563664f * y) * ((-0.625f * xx - 0.625f * yy) + 0.5f);
                const float k12 = (0.373176333f * z) * ((-5.0f * xx - 5.0f * yy) + 2.0f);
                const float k13 = (0.457045799f * x) * ((-5.0f * xx - 5.0f * yy) + 4.0f);
                const float k14 = ((1.44530572f * z) * (x - y)) * (x + y);
                const float k15 = (0.59004359f * x) * (xx - 3.0f * yy);
#pragma unroll
                for (int ch = 0; ch < 3; ++ch) {
                    const float t = (((((k9 * c[27 + ch] + k10 * c[30 + ch]) + k11 * c[33 + ch]) +
                                       k12 * c[36 + ch]) + k13 * c[39 + ch]) + k14 * c[42 + ch]) +
                                    k15 * c[45 + ch];
                    res[ch] = res[ch] + t;
                }
            }
        }
    }
    float4 out;
    out.x = gclamp(res[0] + 0.5f, 0.0f, 1.0f);
    out.y = gclamp(res[1] + 0.5f, 0.0f, 1.0f);
    out.z = gclamp(res[2] + 0.5f, 0.0f, 1.0f);
    out.w = opacity;
    return out;
}

// The coefficients of Gaussian i of SH group `grp` (lane k of the group) in registers, and its
// opacity: one 4-byte (2-byte) load per coefficient and lane, each wave-coalesced (staging the
// group's contiguous block through LDS with 16-byte loads measured slower: 289 vs 227 us at 6M
// Gaussians).
template <int DEG, typename T>
__device__ __forceinline__ void sh_load(const void* shp, int sh_count, int grp, int k,
                                        float (&c)[3 * (DEG + 1) * (DEG + 1)]) {
    constexpr int NC = 3 * (DEG + 1) * (DEG + 1);
    const T* sh = static_cast<const T*>(shp) + (size_t)grp * kShGroup * sh_count + k;
#pragma unroll
    for (int j = 0; j < NC; ++j) c[j] = sh_ld<T>(sh + (size_t)j * kShGroup);
}

template <int DEG, typename T>
__global__ __launch_bounds__(256) void sh_kernel(ShArgs a) {
    constexpr int NC = 3 * (DEG + 1) * (DEG + 1);
    const int i = blockIdx.x * kShGroup + threadIdx.x;
    if (i >= a.count) return;
    const float4 p = a.position[i];
    float c[NC];
    sh_load<DEG, T>(a.sh, a.sh_count, blockIdx.x, threadIdx.x, c);
    float opacity;
    if constexpr (sizeof(T) == 2) {
        opacity = h16<T>(a.packed[i].y >> 16);   // opacity of the precision form
    } else {
        opacity = p.w;
    }
    for (int v = 0; v < a.nviews; ++v)
        a.color[(size_t)v * a.count + i] = sh_color<DEG>(c, p, view_of(v, a.v0, a.views).cam.position, opacity);
}

template <typename T>
static hipError_t launch_sh_t(const ShArgs& a, int degree, hipStream_t s) {
    const dim3 grid((a.count + kShGroup - 1) / kShGroup), block(kShGroup);
    switch (degree) {
        case 0: hipLaunchKernelGGL((sh_kernel<0, T>), grid, block, 0, s, a); break;
        case 1: hipLaunchKernelGGL((sh_kernel<1, T>), grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL((sh_kernel<2, T>), grid, block, 0, s, a); break;
        case 3: hipLaunchKernelGGL((sh_kernel<3, T>), grid, block, 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_sh(const ShArgs& a, int degree, int precision, hipStream_t s) {
    if (a.count <= 0) return hipSuccess;
    switch (precision) {
        case 0: return launch_sh_t<float>(a, degree, s);
        case 1: return launch_sh_t<__half>(a, degree, s);
        case 2: return launch_sh_t<uint16_t>(a, degree, s);
        default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------
// EWA projection: core/gaussian_render.cu:135-215 (quaternion decode :45-48).
// The eigenvector (cos, sin) is written directly instead of going through
// atan2 -> cos/sin (core/gaussian_render.cu:200,212); the two differ by about
// one ulp and the direct form is exactly reproducible on the host.
// ---------------------------------------------------------------------------
// World covariance of one Gaussian (core/gaussian_render.cu:140-154): scale and the packed
// quaternion bits from the FP32 form (scale_rotation) or the precision form (16-bit scale in
// `packed`, quaternion bits in position.w).
template <typename T>
__device__ __forceinline__ m3 world_sigma(const float4& pd, const float4* scale_rotation, const uint2* packed, int i) {
    float4 sr;
    if constexpr (sizeof(T) == 2) {
        const uint2 pk = packed[i];
        sr = make_float4(h16<T>(pk.x & 0xffffu), h16<T>(pk.x >> 16), h16<T>(pk.y & 0xffffu), pd.w);
    } else {
        sr = scale_rotation[i];
    }
    m3 S;
    S.c[0] = v3{sr.x, 0.0f, 0.0f};
    S.c[1] = v3{0.0f, sr.y, 0.0f};
    S.c[2] = v3{0.0f, 0.0f, sr.z};
    const uint32_t q8 = __float_as_uint(sr.w);
    const float rx = ((float)((q8 >> 24u) & 0xffu) / 255.0f) * 2.0f - 1.0f;
    const float ry = ((float)((q8 >> 16u) & 0xffu) / 255.0f) * 2.0f - 1.0f;
    const float rz = ((float)((q8 >> 8u) & 0xffu) / 255.0f) * 2.0f - 1.0f;
    const float rw = ((float)(q8 & 0xffu) / 255.0f) * 2.0f - 1.0f;
    const m3 RS = mul33(quat_to_m3(rw, rx, ry, rz), S);
    return mul33(RS, transpose3(RS));
}

struct ProjOut {
    float cx, cy, cz;       // clip position (-128 when culled)
    float4 ellipse;         // cos, sin, extent x, extent y
    float4 conic;           // (a, b, c, 0) of the inverse 2D covariance
};

// One Gaussian in one view: EWA projection, core/gaussian_render.cu:156-213.
__device__ __forceinline__ ProjOut project_view(const float4& pd, const m3& sigma, const ViewDesc& vd) {
    const CameraArgs& cam = vd.cam;
    const FrameGeom& geom = vd.geom;
    const v4 vp4 = mul44v(cam.view, v4{pd.x, pd.y, pd.z, 1.0f});
    const v3 v = v3{vp4.x, vp4.y, vp4.z};
    const v2 fc = cam.fov_cotangent;
    const v2 dsb = cam.depth_scale_bias;

    const float zRcp = 1.0f / v.z;
    const float zRcpSqr = zRcp * zRcp;
    const float scaleX = -fc.x * zRcp;
    const float scaleY = -fc.y * zRcp;
    const float ftx = fc.x * v.x, fty = fc.y * v.y;
    const float shearX = ftx * zRcpSqr;
    const float shearY = fty * zRcpSqr;
    const float transX = -ftx * zRcp;
    const float transY = -fty * zRcp;

    m4 A;
    A.c[0] = v4{scaleX, 0.0f, 0.0f, 0.0f};
    A.c[1] = v4{0.0f, scaleY, 0.0f, 0.0f};
    A.c[2] = v4{shearX, shearY, dsb.x, 0.0f};
    A.c[3] = v4{transX, transY, dsb.y, 1.0f};

    const m3 VP = mul33(upper3(A), upper3(cam.view));
    m3 cc = mul33(mul33(VP, transpose3(sigma)), transpose3(VP));
    const v4 clip4 = mul44v(A, v4{v.x, v.y, v.z, 1.0f});
    v3 clip = v3{clip4.x, clip4.y, clip4.z};

    cc.c[0].x = cc.c[0].x + geom.trace_bump_x;
    cc.c[1].y = cc.c[1].y + geom.trace_bump_y;

    const float ca = cc.c[0].x, cb = cc.c[1].x, ccc = cc.c[1].y;
    const float det = ca * ccc - cb * cb;
    const float mid = 0.5f * (ca + ccc);
    const float eps = 1e-12f;
    const float radius = sqrtf(gmax(eps, mid * mid - det));
    const float lambda0 = mid + radius;
    const float lambda1 = gmax(0.0f, mid - radius);
    const v2 e0 = normalize2(v2{cb, lambda0 - ca});
    float extx = sqrtf(lambda0) * 3.0f;
    float exty = sqrtf(lambda1) * 3.0f;
    const float invDet = 1.0f / gmax(eps, det);
    const float conx = ccc * invDet, cony = (-cb) * invDet, conz = ca * invDet;

    const float ex = gstep(-1.0f, clip.x) * gstep(clip.x, 1.0f);
    const float ey = gstep(-1.0f, clip.y) * gstep(clip.y, 1.0f);
    const float ez = gstep(-1.0f, clip.z) * gstep(clip.z, 1.0f);
    const float vis = ((ex * ey) * ez) * gstep(0.0f, lambda1);
    clip.x = -128.0f * (1.0f - vis) + clip.x * vis;
    clip.y = -128.0f * (1.0f - vis) + clip.y * vis;
    clip.z = -128.0f * (1.0f - vis) + clip.z * vis;
    extx = extx * vis;
    exty = exty * vis;

    ProjOut o;
    o.cx = clip.x;
    o.cy = clip.y;
    o.cz = clip.z;
    o.ellipse = make_float4(e0.x, e0.y, extx, exty);
    o.conic = make_float4(conx, cony, conz, 0.0f);
    return o;
}

// Whether a projected Gaussian has a non-empty tile rect (it then gets build candidates).
__device__ __forceinline__ bool has_tiles(const ProjOut& o, const FrameGeom& g) {
    const Footprint f = footprint(o.cx, o.cy, o.ellipse.x, o.ellipse.y, o.ellipse.z, o.ellipse.w, g);
    return f.nx > 0 && f.ny > 0;
}

// One lane per Gaussian, every view of the batch.
template <typename T>
__global__ __launch_bounds__(256) void project_kernel(ProjectArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.count) return;
    const float4 pd = a.position[i];
    const m3 sigma = world_sigma<T>(pd, a.scale_rotation, a.packed, i);
    for (int vi = 0; vi < a.nviews; ++vi) {
        const ViewDesc& vd = view_of(vi, a.v0, a.views);
        const size_t o = (size_t)vi * a.count + i;
        const ProjOut po = project_view(pd, sigma, vd);
        a.clip_xy[o] = make_float2(po.cx, po.cy);
        a.clip_z[o] = po.cz;
        a.ellipse[o] = po.ellipse;
        a.conic[o] = po.conic;
    }
}

hipError_t launch_project(const ProjectArgs& a, hipStream_t s) {
    if (a.count <= 0) return hipSuccess;
    const dim3 grid((a.count + 255) / 256), block(256);
    switch (a.precision) {
        case 0: hipLaunchKernelGGL(project_kernel<float>, grid, block, 0, s, a); break;
        case 1: hipLaunchKernelGGL(project_kernel<__half>, grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL(project_kernel<uint16_t>, grid, block, 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Tile-list build.
// Footprint: getAABBRect core/gaussian_render.cu:274-298 and the tile rect of
// :379-389.  Exact test: ellipseRectOverlap :250-272 with
// convertToEllipseCoordinates :234-239 and intersectsUnitCircle :241-248.
// Key: getKey :306-310.  The reference lets atomics decide entry order
// (:485, :498); here entries follow Gaussian index, then row-major candidate
// order, so the stable sort that follows is deterministic.  The record walk
// runs over the compacted visible Gaussians of all views (a wave's 64 records
// may belong to different views); its count pass writes each record's entry
// count and candidate mask and its emit pass writes each record's entries at
// an offset assigned in depth order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t depth_word(float z) {
    return f2u_sat(gclamp((z + 1.0f) * 0.5f, 0.0f, 1.0f) * 4294967296.0f);
}

// q = t / n, r = t % n for 0 <= t < 2^20, 1 <= n < 2^16 (float estimate + one fix)
__device__ __forceinline__ void divmod_small(int t, int n, int& q, int& r) {
    q = (int)((float)t * __builtin_amdgcn_rcpf((float)n));
    r = t - q * n;
    if (r < 0) { --q; r += n; }
    else if (r >= n) { ++q; r -= n; }
}

__global__ void fill_u32_kernel(uint32_t* p, uint32_t value, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

// Wave-wide inclusive scan.
__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int n = __shfl_up(v, d, 64);
        if (lane >= d) v += n;
    }
    return v;
}

// Owner (wave-local record 0..63) of item base + lane of a step, when record `lane` holds items
// [excl, excl + cnt) (excl non-decreasing over the records, base a multiple of 64): every record
// with items in the step marks the position of its first one with the step's tag, a ballot
// collects the marks and each lane takes the nearest mark at or below it -- one LDS write, one
// read and a cross-lane read instead of a binary search of six dependent LDS reads.  s_mark is
// the wave's 64 slots, set to ~0 (no valid tag) before the first step.  Lanes past the step's
// last item get an arbitrary owner.
__device__ __forceinline__ int step_owner(uint32_t* s_mark, int base, int excl, int cnt, int lane) {
    const uint32_t tag = (uint32_t)base;   // low 6 bits free: base is a multiple of 64
    if (cnt > 0 && excl < base + 64 && excl + cnt > base) s_mark[max(excl, base) - base] = tag | (uint32_t)lane;
    wave_sync();
    const uint32_t v = s_mark[lane];
    const uint64_t starts = __ballot((v & ~63u) == tag);
    const uint64_t le = starts & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
    const int slot = le ? 63 - __clzll(le) : 0;
    return __shfl((int)(v & 63u), slot, 64);
}

// Position of the k-th set bit of m (k < popcount(m)), by halving.
__device__ __forceinline__ int select_bit(uint64_t m, int k) {
    int pos = 0;
    uint32_t x = (uint32_t)m;
    int c = __popc(x);
    if (k >= c) {
        k -= c;
        pos = 32;
        x = (uint32_t)(m >> 32);
    }
    c = __popc(x & 0xffffu);
    if (k >= c) { k -= c; pos += 16; x >>= 16; }
    c = __popc(x & 0xffu);
    if (k >= c) { k -= c; pos += 8; x >>= 8; }
    c = __popc(x & 0xfu);
    if (k >= c) { k -= c; pos += 4; x >>= 4; }
    c = __popc(x & 0x3u);
    if (k >= c) { k -= c; pos += 2; x >>= 2; }
    return pos + (k >= (int)(x & 1u) ? 1 : 0);
}

__device__ __forceinline__ uint64_t low_bits(int n) { return n >= 64 ? ~0ull : ((1ull << n) - 1ull); }

// Index-order walk (the stage-level gr_build_tile_list): one wave per
// 64-Gaussian batch of one view.  The wave lays its Gaussians' candidate tiles
// (rect cells, row-major) end to end and tests 64 candidates per step, one per
// lane; the count pass writes the batch's entry count, the emit pass writes
// the entries at the batch's scanned offset in (Gaussian, row, column) order.
template <bool EMIT>
__global__ __launch_bounds__(kBuildThreads) void build_kernel(BuildArgs a) {
    constexpr int W = kBuildThreads / 64;
    __shared__ int s_excl[W][64];
    __shared__ int4 s_rect[W][64];          // x0, y0, nx, depth word
    __shared__ EllipseRec s_ell[W][64];
    __shared__ FastEllipse s_fast[W][64];
    __shared__ uint32_t s_mark[W][64];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int batch = blockIdx.x * W + w;
    if (batch >= a.batches) return;   // wave-uniform; no block barriers below
    const int view = blockIdx.y;
    const ViewDesc& vd = view_of(view, a.v0, a.views);
    const FrameGeom& geom = vd.geom;
    const int first = batch * kBuildBatch;
    int cnt = 0;
    if (first + lane < a.count) {
        const size_t vo = (size_t)view * a.count + first + lane;
        const float2 c = a.clip_xy[vo];
        const float4 e = a.ellipse[vo];
        const Footprint f = footprint(c.x, c.y, e.x, e.y, e.z, e.w, geom);
        // the reference keeps max(0, w*h); a rect with w <= 0 or h <= 0 only
        // arises from NaN corners and never passes the exact test
        cnt = (f.nx > 0 && f.ny > 0) ? f.nx * f.ny : 0;
        if (cnt > 0) {
            s_rect[w][lane] = make_int4(f.x0, f.y0, f.nx, EMIT ? (int)depth_word(a.clip_z[vo]) : 0);
            const EllipseRec er{c.x, c.y, e.x, e.y, e.z, e.w, 1.0f / e.z, 1.0f / e.w};
            s_ell[w][lane] = er;
            s_fast[w][lane] = fast_ellipse(er, f, geom.tile_clip_x, geom.tile_clip_y);
        }
    }
    const int incl = wave_incl_scan(cnt, lane);
    const int total = __shfl(incl, 63, 64);
    const int excl = incl - cnt;
    s_excl[w][lane] = excl;
    s_mark[w][lane] = ~0u;
    wave_sync();
    const size_t bslot = (size_t)view * a.batches + batch;
    const uint32_t out_base = EMIT ? a.batch_counts[bslot] : 0u;
    uint32_t running = 0;
    for (int base = 0; base < total; base += 64) {
        const int cc = base + lane;
        bool pass = false;
        const int o = step_owner(s_mark[w], base, excl, cnt, lane);
        uint32_t tile = 0;
        if (cc < total) {
            const int4 rc = s_rect[w][o];
            int q, r;
            divmod_small(cc - s_excl[w][o], rc.z, q, r);
            pass = tile_overlap_fast(rc.x + r, rc.y + q, s_ell[w][o], s_fast[w][o], geom);
            tile = (uint32_t)(vd.tile_base + (rc.y + q) * geom.tiles_x + rc.x + r);
        }
        const uint64_t m = __ballot(pass);
        if (EMIT && pass) {
            const uint32_t p = out_base + running + (uint32_t)__popcll(m & lanemask_lt(lane));
            if (p < (uint32_t)a.capacity) {
                a.keys[p] = ((uint64_t)tile << 32) | (uint32_t)s_rect[w][o].w;
                a.values[p] = first + o;
            }
        }
        running += (uint32_t)__popcll(m);
        wave_sync();
    }
    if (!EMIT && lane == 0) a.batch_counts[bslot] = running;
}

// Record walk, count pass: one wave per 64 consecutive records in index
// order (which spreads small and large footprints evenly over the waves; a
// wave's records may belong to different views).  Per record it writes the
// exact entry count and a 64-bit mask whose bit l says whether candidate l
// (row-major in the footprint) passed the exact test -- the whole answer for
// footprints of at most 64 tiles, so the emit pass needs no tests for them.
// Each record lane gathers its own count and mask from each step's ballot.
// (Walking the records straight from the projection's index-order arrays,
// without compacting them first, measured slower: 62 % of their lanes are
// visible, so the waves are more and emptier.)
__global__ __launch_bounds__(kBuildThreads) void record_count_kernel(BuildArgs a) {
    constexpr int W = kBuildThreads / 64;
    __shared__ int s_excl[W][64];
    __shared__ int4 s_rect[W][64];       // x0, y0, nx, candidates
    __shared__ EllipseRec s_ell[W][64];
    __shared__ float4 s_fast[W][64];     // kv, kh, lo of the bounded test; tile width (clip units)
    __shared__ float s_tcy[W][64];       // tile height (clip units) of the record's view
    __shared__ uint32_t s_mark[W][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int limit = (int)min(*a.record_total, (uint32_t)a.nviews * (uint32_t)a.count);
    const int first = (blockIdx.x * W + w) * kBuildBatch;
    if (first >= limit) return;   // wave-uniform; no block barriers below
    const int pos = first + lane;
    int cnt = 0;
    uint32_t slot = 0u, dword = 0u, pg = 0u, corner = 0u;
    if (pos < limit) {
        slot = a.record_slots[pos];
        const float4 r0 = a.records[2 * (size_t)slot];
        const float4 r1 = a.records[2 * (size_t)slot + 1];
        dword = __float_as_uint(r1.z);
        pg = __float_as_uint(r1.w);
        const int view = (int)pg / a.count;
        const FrameGeom& g = view_of(view, a.v0, a.views).geom;
        const Footprint f = footprint(r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, g);
        cnt = (f.nx > 0 && f.ny > 0) ? f.nx * f.ny : 0;
        if (cnt > 0 && cnt <= 64)
            corner = (uint32_t)(f.y0 * g.tiles_x + f.x0) | ((uint32_t)(f.nx - 1) << kEmitCornerBits);
        s_rect[w][lane] = make_int4(f.x0, f.y0, f.nx, cnt);
        const EllipseRec er{r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, 1.0f / r1.x, 1.0f / r1.y};
        s_ell[w][lane] = er;
        const FastEllipse fe = fast_ellipse(er, f, g.tile_clip_x, g.tile_clip_y);
        s_fast[w][lane] = make_float4(fe.kv, fe.kh, fe.lo, g.tile_clip_x);
        s_tcy[w][lane] = g.tile_clip_y;
    }
    s_mark[w][lane] = ~0u;
    const int incl = wave_incl_scan(cnt, lane);
    const int total = __shfl(incl, 63, 64);
    const int excl = incl - cnt;
    s_excl[w][lane] = excl;
    // footprints over 64 candidates take cnt words each of the wave's slice of a.side (slices of
    // equal size, one per wave of the frame's records: no atomics); those that fit list their
    // passing tiles there, in entry order
    uint32_t side_at = ~0u;
    if (a.side) {
        const int64_t slice = a.side_words / (((int64_t)limit + kBuildBatch - 1) / kBuildBatch);
        const int want = cnt > 64 ? cnt : 0;
        const int wexcl = wave_incl_scan(want, lane) - want;
        if (want > 0 && (int64_t)wexcl + want <= slice)
            side_at = (uint32_t)((int64_t)(first / kBuildBatch) * slice + wexcl);
    }
    const bool any_listed = __ballot(side_at != ~0u) != 0ull;   // wave-uniform
    wave_sync();
    // this lane's record: its entries and the mask of its passing candidates 0..63, gathered
    // from each step's ballot
    uint32_t my_cnt = 0u;
    uint64_t my_mask = 0ull;
    for (int base = 0; base < total; base += 64) {
        const int cc = base + lane;
        bool pass = false;
        const int o = step_owner(s_mark[w], base, excl, cnt, lane);
        // a listed record's list word and entries found so far, from its own lane (every lane
        // active for the shuffles)
        uint32_t o_side = ~0u, o_done = 0u;
        if (any_listed) {
            o_side = __shfl(side_at, o, 64);
            o_done = __shfl(my_cnt, o, 64);
        }
        uint32_t tile = 0u;
        if (cc < total) {
            const int l = cc - s_excl[w][o];
            const int4 rc = s_rect[w][o];
            int q, r;
            divmod_small(l, rc.z, q, r);
            if (any_listed) tile = (uint32_t)((rc.y + q) * a.v0.geom.tiles_x + rc.x + r);   // one view
            const float4 sf = s_fast[w][o];
            const EllipseRec& er = s_ell[w][o];
            FrameGeom og;
            og.tile_clip_x = sf.w;
            og.tile_clip_y = s_tcy[w][o];
            pass = tile_overlap_fast(rc.x + r, rc.y + q, er, fast_ellipse_from(er, sf.x, sf.y, sf.z), og);
        }
        const uint64_t m = __ballot(pass);
        if (any_listed && pass && o_side != ~0u) {
            // the record's lanes in this step start at lane max(0, its first candidate - base)
            const int lo2 = max(0, s_excl[w][o] - base);
            const uint64_t mo = m & lanemask_lt(lane) & ~lanemask_lt(lo2);
            a.side[o_side + o_done + (uint32_t)__popcll(mo)] = tile;
        }
        if (cnt > 0 && excl < base + 64 && excl + cnt > base) {
            const int st = max(excl, base);
            const uint64_t bits = (m >> (st - base)) & low_bits(min(excl + cnt, base + 64) - st);
            my_cnt += (uint32_t)__popcll(bits);
            if (st - excl < 64) my_mask |= bits << (st - excl);
        }
    }
    if (pos < limit) {
        // the depth sort's pair: depth word, record index | count (saturated) in the bits above it
        const uint32_t esc = 0xffffffffu >> a.index_bits;
        a.depth_keys[pos] = dword;
        a.depth_words[pos] = (uint32_t)pos | (min(my_cnt, esc) << a.index_bits);
        a.emit_records[pos] = cnt <= 64      ? make_uint4((uint32_t)my_mask, (uint32_t)(my_mask >> 32), pg, corner)
                              : side_at != ~0u ? make_uint4(side_at, my_cnt, pg, kEmitBig | kEmitSide)
                                               : make_uint4(slot, my_cnt, pg, kEmitBig);
    }
}

// Record walk, emit pass, in depth order.  The list is cut into windows of kEmitWindow entries;
// a wave writes one window at a time (windows taken grid-stride, one tile-digit histogram flush
// per workgroup).  windows[w] is the depth-order position of the record holding the window's
// first entry; from there the wave reads the records' packed words and output offsets
// contiguously, 64 positions at a time, fetches their emit records, and writes the window's
// entries in order: window entry e belongs to the last record starting at or before it and is
// that record's entry k = e - (its start).  Footprints of at most 64 candidates give entry k as
// candidate select_bit(mask, k) (no tests, no footprint); larger ones whose passing tiles the
// count pass listed read entry k from the list; the others fetch their build record, run the
// overlap tests again and keep the passing candidates whose entry index falls in the window.  Entries are (Gaussian index << 32 | tile id); their positions follow depth order, so
// the depth word of the key is carried by the position and the stable tile sort that follows
// needs the tile bits alone.  Every window holds the same number of entries, so the work is
// spread evenly whatever the footprint sizes, and the list is written front to back.
#ifndef GR_EMIT_WPE
#define GR_EMIT_WPE 4
#endif
__global__ __launch_bounds__(kBuildThreads) __attribute__((amdgpu_waves_per_eu(GR_EMIT_WPE))) void record_emit_kernel(
    BuildArgs a) {
    constexpr int W = kBuildThreads / 64;
    __shared__ int s_rel[W][64];         // record start relative to the window (negative: earlier)
    __shared__ int s_excl[W][64];        // first candidate of a large footprint in the step order
    __shared__ int4 s_rect[W][64];       // global tile of the rect corner, nx, tiles_x, -
    __shared__ uint64_t s_mask[W][64];
    __shared__ int s_gi[W][64];
    __shared__ uint32_t s_xy[W][64];     // rect corner in view tiles, x | y << 16 (large footprints)
    __shared__ EllipseRec s_ell[W][64];  // large footprints only
    __shared__ int s_view[W][64];
    __shared__ uint32_t s_done[W][64];
    __shared__ int s_cand[W][64];        // candidates of a large footprint (0: small footprint)
    __shared__ int2 s_keep[W][64];       // entry indices [lo, hi) of a large footprint in the window
    __shared__ float4 s_fast[W][64];     // kv, kh, lo of the bounded test (large footprints)
    __shared__ uint32_t s_mark[W][64];
    __shared__ uint32_t s_hist[3][256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool hist = a.tile_hist != nullptr;
    const int db = a.tile_digit_bits;
    const uint32_t dm = (1u << db) - 1u;
    if (hist)
        for (int i = threadIdx.x; i < 3 * 256; i += kBuildThreads) (&s_hist[0][0])[i] = 0u;
    __syncthreads();
    const uint32_t nrec = min(*a.record_total, (uint32_t)a.nviews * (uint32_t)a.count);
    const uint32_t kend = min(*a.entry_total, (uint32_t)a.capacity);
    const int64_t nwin = ((int64_t)kend + kEmitWindow - 1) / kEmitWindow;
    const int ib = a.index_bits;
    const uint32_t imask = (1u << ib) - 1u;
    const uint32_t esc = 0xffffffffu >> ib;
    auto put = [&](uint32_t p, uint32_t tile, int o) {
        a.entries[p] = ((uint64_t)(uint32_t)s_gi[w][o] << 32) | tile;
        if (hist)
            for (int k = 0; k < a.tile_passes; ++k) atomicAdd(&s_hist[k][(tile >> (db * k)) & dm], 1u);
    };
    // a lane's depth-order position q + lane: its output offset and packed word, then (when it
    // starts before kmax, the end of its window) its emit record
    auto load_words = [&](uint32_t q, uint32_t& p, uint32_t& word) {
        const uint32_t qi = q + (uint32_t)lane;
        p = qi < nrec ? a.depth_prefix[qi] : 0xffffffffu;
        word = qi < nrec ? a.sorted_words[qi] : 0u;
    };
    auto load_rec = [&](uint32_t p, uint32_t word, uint32_t kmax) {
        return p < kmax ? a.emit_records[word & imask] : make_uint4(0u, 0u, 0u, 0u);
    };
    auto win_end = [&](int64_t wv) { return min((uint32_t)wv * (uint32_t)kEmitWindow + (uint32_t)kEmitWindow, kend); };
    // loads run two windows ahead: while window i is written, the emit records of window i + 1's
    // first 64 positions, the offsets and words of window i + 2's and the first position of
    // window i + 3 are in flight (each load has a whole window's work to arrive)
    const int64_t wstride = (int64_t)gridDim.x * W;
    int64_t win = (int64_t)blockIdx.x * W + w;
    uint32_t q0 = 0u, p0 = 0xffffffffu, word0 = 0u;   // window i
    uint4 er0 = make_uint4(0u, 0u, 0u, 0u);
    uint32_t q1 = 0u, p1 = 0xffffffffu, word1 = 0u;   // window i + 1
    uint32_t q2 = 0u;                                 // window i + 2
    if (win < nwin) {
        q0 = a.windows[win];
        load_words(q0, p0, word0);
        er0 = load_rec(p0, word0, win_end(win));
    }
    if (win + wstride < nwin) {
        q1 = a.windows[win + wstride];
        load_words(q1, p1, word1);
    }
    if (win + 2 * wstride < nwin) q2 = a.windows[win + 2 * wstride];
    for (; win < nwin; win += wstride) {
        const uint32_t k0 = (uint32_t)win * (uint32_t)kEmitWindow;
        const int wlen = (int)(win_end(win) - k0);
        uint32_t q = q0, p = p0, word = word0;
        uint4 er = er0;
        // advance the pipeline
        uint4 er1 = make_uint4(0u, 0u, 0u, 0u);
        uint32_t p2 = 0xffffffffu, word2 = 0u, q3 = 0u;
        if (win + wstride < nwin) er1 = load_rec(p1, word1, win_end(win + wstride));
        if (win + 2 * wstride < nwin) load_words(q2, p2, word2);
        if (win + 3 * wstride < nwin) q3 = a.windows[win + 3 * wstride];
        for (bool first_chunk = true;; first_chunk = false) {
            if (!first_chunk) {
                load_words(q, p, word);
                er = load_rec(p, word, k0 + (uint32_t)wlen);
            }
            // the 64 positions from q; those starting inside the window take part
            const bool in = p < k0 + (uint32_t)wlen;
            int rel = 0, lo = 0, hi = 0, big = 0;
            if (in) {
                const bool small = !(er.w & kEmitBig);
                const uint64_t mask = (uint64_t)er.x | ((uint64_t)er.y << 32);
                const uint32_t pc = word >> ib;
                const int cnt = (int)(pc != esc ? pc : (small ? (uint32_t)__popcll(mask) : er.y));
                rel = (int)(p - k0);   // negative for the record that began in an earlier window
                lo = max(rel, 0);
                hi = (int)min((int64_t)rel + cnt, (int64_t)wlen);
                const int pg = (int)er.z;
                const int view = pg / a.count;
                const ViewDesc& vd = view_of(view, a.v0, a.views);
                s_gi[w][lane] = pg - view * a.count;
                s_rel[w][lane] = rel;
                if (small) {
                    s_mask[w][lane] = mask;
                    const int corner = (int)(er.w & ((1u << kEmitCornerBits) - 1u));
                    s_rect[w][lane] = make_int4(vd.tile_base + corner, (int)(er.w >> kEmitCornerBits) + 1,
                                                vd.geom.tiles_x, 0);
                    s_cand[w][lane] = 0;
                } else if (er.w & kEmitSide) {
                    // listed by the count pass: entry k is the tile at side[x + k]
                    s_rect[w][lane] = make_int4(-1, 0, 0, (int)er.x);
                    s_cand[w][lane] = 0;
                } else if (hi > lo) {
                    const float4 r0 = a.records[2 * (size_t)er.x];
                    const float4 r1 = a.records[2 * (size_t)er.x + 1];
                    const Footprint f = footprint(r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, vd.geom);
                    big = f.nx * f.ny;
                    s_cand[w][lane] = big;
                    s_keep[w][lane] = make_int2(lo - rel, hi - rel);
                    s_rect[w][lane] =
                        make_int4(vd.tile_base + f.y0 * vd.geom.tiles_x + f.x0, f.nx, vd.geom.tiles_x, 0);
                    s_xy[w][lane] = (uint32_t)f.x0 | ((uint32_t)f.y0 << 16);
                    s_view[w][lane] = view;
                    const EllipseRec el{r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, 1.0f / r1.x, 1.0f / r1.y};
                    s_ell[w][lane] = el;
                    const FastEllipse fe = fast_ellipse(el, f, vd.geom.tile_clip_x, vd.geom.tile_clip_y);
                    s_fast[w][lane] = make_float4(fe.kv, fe.kh, fe.lo, 0.0f);
                } else {
                    s_cand[w][lane] = 0;
                }
            }
            // the window entries [first, last) these records hold, one per lane and step; entries of
            // large footprints are left to the candidate loop below
            const int span = hi > lo ? hi - lo : 0;
            const uint64_t held = __ballot(span > 0);
            const int first = held ? __shfl(lo, __builtin_ctzll(held), 64) : 0;
            const int last = held ? __shfl(hi, 63 - __clzll(held), 64) : 0;
            s_mark[w][lane] = ~0u;
            wave_sync();
            for (int base = first & ~63; base < last; base += 64) {
                const int e = base + lane;
                const int o = step_owner(s_mark[w], base, lo, span, lane);
                if (e >= first && e < last && s_cand[w][o] == 0) {
                    const int k = e - s_rel[w][o];
                    const int4 rc = s_rect[w][o];
                    uint32_t tile;
                    if (rc.x < 0) {
                        tile = a.side[(uint32_t)rc.w + (uint32_t)k];
                    } else {
                        const int l = select_bit(s_mask[w][o], k);
                        int qq, rr;
                        divmod_small(l, rc.y, qq, rr);
                        tile = (uint32_t)(rc.x + qq * rc.z + rr);
                    }
                    put(k0 + (uint32_t)e, tile, o);
                }
                wave_sync();
            }
            // candidates of the large footprints, laid end to end
            int total = 0, excl = 0;
            if (__ballot(big > 0)) {   // wave-uniform; most chunks have none
                const int incl = wave_incl_scan(big, lane);
                total = __shfl(incl, 63, 64);
                excl = incl - big;
                wave_sync();
                s_excl[w][lane] = excl;
                s_done[w][lane] = 0u;
                s_mark[w][lane] = ~0u;
                wave_sync();
            }
            for (int base = 0; base < total; base += 64) {
                const int cc = base + lane;
                bool pass = false;
                int l = 0;
                const int o = step_owner(s_mark[w], base, excl, big, lane);
                uint32_t tile = 0u;
                if (cc < total) {
                    l = cc - s_excl[w][o];
                    const int4 rc = s_rect[w][o];
                    const uint32_t xy = s_xy[w][o];
                    int qq, rr;
                    divmod_small(l, rc.y, qq, rr);
                    const FrameGeom& vg = view_of(s_view[w][o], a.v0, a.views).geom;
                    FrameGeom og;
                    og.tile_clip_x = vg.tile_clip_x;
                    og.tile_clip_y = vg.tile_clip_y;
                    const float4 sf = s_fast[w][o];
                    const EllipseRec& el = s_ell[w][o];
                    pass = tile_overlap_fast((int)(xy & 0xffffu) + rr, (int)(xy >> 16) + qq, el,
                                             fast_ellipse_from(el, sf.x, sf.y, sf.z), og);
                    tile = (uint32_t)(rc.x + qq * rc.z + rr);
                }
                const uint64_t m = __ballot(pass);
                if (pass) {
                    // lanes of this record in this step start at lane max(0, excl - base)
                    const int lo2 = max(0, s_excl[w][o] - base);
                    const uint64_t before = m & lanemask_lt(lane) & ~lanemask_lt(lo2);
                    const int k = (int)s_done[w][o] + __popcll(before);   // the record's entry index
                    const int2 keep = s_keep[w][o];
                    if (k >= keep.x && k < keep.y) put(k0 + (uint32_t)(s_rel[w][o] + k), tile, o);
                }
                wave_sync();
                if (cc < total && (l == 0 || lane == 0)) {
                    const int len = min(s_excl[w][o] + s_cand[w][o], base + 64) - cc;
                    s_done[w][o] += (uint32_t)__popcll((m >> lane) & low_bits(len));
                }
                wave_sync();
            }
            wave_sync();
            // more records when all 64 took part and the last one ends inside the window
            if (__ballot(in) != ~0ull) break;
            if (__shfl(hi, 63, 64) >= wlen) break;
            q += 64u;
        }
        q0 = q1;
        p0 = p1;
        word0 = word1;
        er0 = er1;
        q1 = q2;
        p1 = p2;
        word1 = word2;
        q2 = q3;
    }
    if (hist) {
        __syncthreads();
        for (int k = 0; k < a.tile_passes; ++k) {
            const uint32_t v = s_hist[k][threadIdx.x];
            if (v) atomicAdd(&a.tile_hist[(blockIdx.x % kSortHistCopies) * kSortHistCopyStride + k * 256 + threadIdx.x], v);
        }
    }
}

static dim3 build_grid(const BuildArgs& a) {
    constexpr int W = kBuildThreads / 64;
    if (a.records) return dim3((unsigned)((a.batches + W - 1) / W));
    return dim3((a.batches + W - 1) / W, a.nviews);
}

hipError_t launch_build_count(const BuildArgs& a, hipStream_t s) {
    if (a.count <= 0 || a.nviews <= 0) return hipSuccess;
    if (a.records) {
        if (!a.record_slots || !a.record_total || !a.depth_keys || !a.depth_words || !a.emit_records ||
            a.index_bits < 1 || a.index_bits > 31 ||
            ((int64_t)1 << a.index_bits) < (int64_t)a.nviews * a.count ||
            (a.side && (a.nviews != 1 || a.side_words <= 0 || a.side_words > (int64_t)UINT32_MAX)))
            return hipErrorInvalidValue;
        hipLaunchKernelGGL(record_count_kernel, build_grid(a), dim3(kBuildThreads), 0, s, a);
    } else {
        hipLaunchKernelGGL(build_kernel<false>, build_grid(a), dim3(kBuildThreads), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_build_emit(const BuildArgs& a, hipStream_t s) {
    if (a.count <= 0 || a.nviews <= 0) return hipSuccess;
    if (a.records) {
        if (!a.record_total || !a.entry_total || !a.sorted_words || !a.depth_prefix || !a.windows ||
            !a.emit_records || !a.records || !a.entries || a.index_bits < 1 || a.index_bits > 31 ||
            a.tile_passes > 3 || (a.tile_hist && (a.tile_digit_bits < 1 || a.tile_digit_bits > 8)))
            return hipErrorInvalidValue;
        constexpr int W = kBuildThreads / 64;
        const int64_t wins = emit_window_words(a.capacity);
        const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>((wins + W - 1) / W, kEmitBlocks));
        hipLaunchKernelGGL(record_emit_kernel, dim3(blocks), dim3(kBuildThreads), 0, s, a);
    } else {
        hipLaunchKernelGGL(build_kernel<true>, build_grid(a), dim3(kBuildThreads), 0, s, a);
    }
    return hipGetLastError();
}

// Non-temporal stores for the front end's outputs (colour, clip xy / z, conic, build records):
// each is read once, by a later kernel, long after it is written (6M/1080p frame 852 -> 861
// frames/s against plain stores, same box).
typedef float nt_f4 __attribute__((ext_vector_type(4)));
typedef float nt_f2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st_nt(float4* p, const float4& v) {
    const nt_f4 n = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(n, reinterpret_cast<nt_f4*>(p));
}
__device__ __forceinline__ void st_nt(float2* p, const float2& v) {
    const nt_f2 n = {v.x, v.y};
    __builtin_nontemporal_store(n, reinterpret_cast<nt_f2*>(p));
}
__device__ __forceinline__ void st_nt(float* p, float v) { __builtin_nontemporal_store(v, p); }

static unsigned grid_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

// ---------------------------------------------------------------------------
// Front end: SH colour + projection + compaction into build records, one pass (FrontArgs).
// ---------------------------------------------------------------------------
// The frame's fills (FrontArgs::fills), grid-stride over the front end's lanes.
__device__ __forceinline__ void front_fills(const FrontArgs& a, int64_t gid, int64_t lanes) {
    for (int f = 0; f < a.nfills; ++f) {
        const FillSpan sp = a.fills[f];
        for (int64_t j = gid; j < sp.n; j += lanes) sp.p[j] = sp.value;
    }
}
__global__ __launch_bounds__(kFrontThreads) void front_fills_kernel(FrontArgs a) {
    front_fills(a, (int64_t)blockIdx.x * kFrontThreads + threadIdx.x, (int64_t)gridDim.x * kFrontThreads);
}

// Waves per SIMD the compiler is asked to fit into its register budget (0: its own choice);
// a build-time override for sweeps.
#ifndef GR_FRONT_WPE
#define GR_FRONT_WPE 0
#endif
template <int DEG, typename T>
__global__ __launch_bounds__(kFrontThreads) __attribute__((amdgpu_waves_per_eu(GR_FRONT_WPE))) void front_kernel(
    FrontArgs a) {
    constexpr int NC = 3 * (DEG + 1) * (DEG + 1);
    constexpr int WV = kFrontThreads / 64;
    __shared__ uint64_t s_bits[kFrontMaxViews][WV];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int blk = blockIdx.x;
    const int i = blk * kFrontThreads + t;
    front_fills(a, (int64_t)i, (int64_t)gridDim.x * kFrontThreads);
    const bool in = i < a.count;
    float4 pd = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (in) pd = a.position[i];
    // colours of every view (the coefficients are dead before the projections)
    if (in && a.degree > 0) {
        float c[NC];
        sh_load<DEG, T>(a.sh, a.sh_count, blk, t, c);
        float opacity;
        if constexpr (sizeof(T) == 2) {
            opacity = h16<T>(a.packed[i].y >> 16);
        } else {
            opacity = pd.w;
        }
        for (int v = 0; v < a.nviews; ++v)
            st_nt(&a.color[(size_t)v * a.count + i], sh_color<DEG>(c, pd, view_of(v, a.v0, a.views).cam.position, opacity));
    }
    // projections; the record of view 0 stays in registers
    float4 ra0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), rb0 = ra0;
    {
        m3 sigma{};
        if (in) sigma = world_sigma<T>(pd, a.scale_rotation, a.packed, i);
        for (int v = 0; v < a.nviews; ++v) {
            const ViewDesc& vd = view_of(v, a.v0, a.views);
            bool vis = false;
            if (in) {
                const size_t o = (size_t)v * a.count + i;
                const ProjOut po = project_view(pd, sigma, vd);
                st_nt(&a.clip_xy[o], make_float2(po.cx, po.cy));
                if (a.clip_z) st_nt(&a.clip_z[o], po.cz);
                if (a.ellipse) st_nt(&a.ellipse[o], po.ellipse);
                st_nt(&a.conic[o], po.conic);
                vis = has_tiles(po, vd.geom);
                if (v == 0) {
                    ra0 = make_float4(po.cx, po.cy, po.ellipse.x, po.ellipse.y);
                    rb0 = make_float4(po.ellipse.z, po.ellipse.w, __uint_as_float(depth_word(po.cz)),
                                      __int_as_float(i));
                }
            }
            const uint64_t m = __ballot(vis);
            if (lane == 0) s_bits[v][w] = m;
        }
    }
    __syncthreads();
    // the block's records, (view, index) order, into its slice of the staging array
    float4* slice = a.records + 2 * (size_t)blk * kFrontThreads * a.nviews;
    uint32_t base = 0u;
    for (int v = 0; v < a.nviews; ++v) {
        uint32_t before = 0u, vtot = 0u;
#pragma unroll
        for (int k = 0; k < WV; ++k) {
            const uint32_t c = (uint32_t)__popcll(s_bits[v][k]);
            if (k < w) before += c;
            vtot += c;
        }
        const uint64_t m = s_bits[v][w];
        if ((m >> lane) & 1ull) {
            const uint32_t r = base + before + (uint32_t)__popcll(m & lanemask_lt(lane));
            float4 ra = ra0, rb = rb0;
            if (v > 0) {
                // the projection of view v again (the same arithmetic, so the same bits), rather
                // than a store and reload of its ellipse and depth per Gaussian and view
                const ProjOut po = project_view(pd, world_sigma<T>(pd, a.scale_rotation, a.packed, i),
                                                view_of(v, a.v0, a.views));
                ra = make_float4(po.cx, po.cy, po.ellipse.x, po.ellipse.y);
                rb = make_float4(po.ellipse.z, po.ellipse.w, __uint_as_float(depth_word(po.cz)),
                                 __int_as_float(v * a.count + i));
            }
            st_nt(&slice[2 * (size_t)r], ra);
            st_nt(&slice[2 * (size_t)r + 1], rb);
        }
        base += vtot;
    }
    if (t == 0) a.block_counts[blk] = base;
}

template <typename T>
static hipError_t launch_front_t(const FrontArgs& a, unsigned blocks, hipStream_t s) {
    switch (a.degree) {
        case 0: hipLaunchKernelGGL((front_kernel<0, T>), dim3(blocks), dim3(kFrontThreads), 0, s, a); break;
        case 1: hipLaunchKernelGGL((front_kernel<1, T>), dim3(blocks), dim3(kFrontThreads), 0, s, a); break;
        case 2: hipLaunchKernelGGL((front_kernel<2, T>), dim3(blocks), dim3(kFrontThreads), 0, s, a); break;
        case 3: hipLaunchKernelGGL((front_kernel<3, T>), dim3(blocks), dim3(kFrontThreads), 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_front(const FrontArgs& a, hipStream_t s) {
    if (a.nfills < 0 || a.nfills > kFrontFills) return hipErrorInvalidValue;
    for (int f = 0; f < a.nfills; ++f)
        if (a.fills[f].n < 0 || (a.fills[f].n > 0 && !a.fills[f].p)) return hipErrorInvalidValue;
    if (a.count <= 0) {
        // no Gaussians: the fills alone
        if (a.nfills > 0) {
            hipLaunchKernelGGL(front_fills_kernel, dim3(256), dim3(kFrontThreads), 0, s, a);
            return hipGetLastError();
        }
        return hipSuccess;
    }
    if (a.nviews < 1 || a.nviews > kFrontMaxViews || !a.records || !a.block_counts || !a.clip_xy || !a.conic ||
        (a.degree > 0 && (!a.sh || !a.color)))
        return hipErrorInvalidValue;
    const unsigned blocks = (unsigned)front_blocks(a.count);
    switch (a.precision) {
        case 0: return launch_front_t<float>(a, blocks, s);
        case 1: return launch_front_t<__half>(a, blocks, s);
        case 2: return launch_front_t<uint16_t>(a, blocks, s);
        default: return hipErrorInvalidValue;
    }
}

// One wave per front-end block: its records' global numbers -> their staging slots.
__global__ __launch_bounds__(256) void record_slots_kernel(const uint32_t* offsets, const uint32_t* total,
                                                           int32_t blocks, uint32_t region, uint32_t* slots) {
    const int b = (int)((blockIdx.x * 256u + threadIdx.x) >> 6), lane = threadIdx.x & 63;
    if (b >= blocks) return;
    const uint32_t o = offsets[b];
    const uint32_t e = b + 1 < blocks ? offsets[b + 1] : *total;
    const uint32_t slot = (uint32_t)b * region;
    for (uint32_t r = (uint32_t)lane; o + r < e; r += 64u) slots[o + r] = slot + r;
}

hipError_t launch_record_slots(const uint32_t* offsets, const uint32_t* total, int32_t blocks, int32_t region,
                               uint32_t* slots, hipStream_t s) {
    if (blocks <= 0) return hipSuccess;
    if (!offsets || !total || !slots || region <= 0 || (int64_t)blocks * region > (int64_t)UINT32_MAX)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(record_slots_kernel, dim3((unsigned)((blocks + 3) / 4)), dim3(256), 0, s, offsets, total,
                       blocks, (uint32_t)region, slots);
    return hipGetLastError();
}

// Exclusive scan of the per-batch counts, in place, in two launches (reduce
// per 2048-element block; local scan + block offset, each block summing the
// block sums before its own); the grand total (saturated at 2^32-1) goes to
// *total.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ uint32_t block_sum_256(uint32_t v, uint32_t* s_w) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_down(v, d, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
    __syncthreads();
    return s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// Packed record words (the low words of the depth pairs): record index in
// bits [0, ib), entry count in bits [ib, 32), saturated at 2^(32-ib) - 1, in
// which case the record's emit record gives it (1 <= ib <= 31).
__device__ __forceinline__ uint32_t packed_count(uint32_t v, int ib, const uint4* erec) {
    const uint32_t c = v >> ib;
    if (c != (0xffffffffu >> ib)) return c;
    const uint4 e = erec[v & ((1u << ib) - 1u)];
    return (e.w & kEmitBig) ? e.y : (uint32_t)__popcll((uint64_t)e.x | ((uint64_t)e.y << 32));
}

__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const uint32_t* data, int32_t n, uint32_t* sums) {
    __shared__ uint32_t s_w[4];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = base + k * kScanThreads + threadIdx.x;
        if (i < n) v += data[i];
    }
    const uint32_t tot = block_sum_256(v, s_w);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// The counts of the first *valid packed record words.
__global__ __launch_bounds__(kScanThreads) void scan_reduce_packed_kernel(const uint32_t* words, int32_t n,
                                                                          uint32_t* sums, const uint32_t* valid,
                                                                          const uint4* erec, int ib) {
    __shared__ uint32_t s_w[4];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    const int64_t nv = (int64_t)min(*valid, (uint32_t)n);
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = base + k * kScanThreads + threadIdx.x;
        if (i < nv) v += packed_count(words[i], ib, erec);
    }
    const uint32_t tot = block_sum_256(v, s_w);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// The offset of this block's tile: the sum of the reduce pass's per-block sums before it,
// saturated to 32 bits, summed by the block itself (so that no launch sits between the two
// passes); the last block also writes the total (saturated).
__device__ __forceinline__ uint32_t block_offset_256(const uint32_t* sums, uint32_t* total, uint64_t* s_w64) {
    const int b = blockIdx.x;
    const bool last = b + 1 == (int)gridDim.x;
    const int upto = last ? b + 1 : b;
    unsigned long long v = 0;
    for (int i = threadIdx.x; i < upto; i += kScanThreads) v += sums[i];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_down(v, d, 64);
    if ((threadIdx.x & 63) == 0) s_w64[threadIdx.x >> 6] = v;
    __syncthreads();
    const uint64_t all = s_w64[0] + s_w64[1] + s_w64[2] + s_w64[3];
    const uint64_t excl = last ? all - sums[b] : all;
    if (last && threadIdx.x == 0) *total = (uint32_t)(all > 0xffffffffull ? 0xffffffffull : all);
    return (uint32_t)(excl > 0xffffffffull ? 0xffffffffull : excl);
}

// Local scan of kScanItems values per lane plus the block's offset from the
// scanned block sums; returns this lane's first exclusive prefix.
__device__ __forceinline__ uint32_t block_excl_256(uint32_t sum, uint32_t block_off, uint32_t* s_w) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t incl = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(incl, d, 64);
        if (lane >= d) incl += o;
    }
    if (lane == 63) s_w[w] = incl;
    __syncthreads();
    uint32_t run = block_off + incl - sum;
    for (int k = 0; k < w; ++k) run += s_w[k];
    return run;
}

__global__ __launch_bounds__(kScanThreads) void scan_down_kernel(uint32_t* data, int32_t n, const uint32_t* sums,
                                                                 uint32_t* total) {
    __shared__ uint32_t s_w[4];
    __shared__ uint64_t s_w64[4];
    const int64_t my = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (my + k < n) ? data[my + k] : 0u;
        sum += v[k];
    }
    uint32_t run = block_excl_256(sum, block_offset_256(sums, total, s_w64), s_w);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (my + k < n) data[my + k] = run;
        run += v[k];
    }
}

// The exclusive prefix of packed word j (j < *valid) goes to prefix[j]; every window start
// w * kEmitWindow (w < nwin) inside word j's entries [prefix, prefix + count) gets windows[w] = j.
__global__ __launch_bounds__(kScanThreads) void scan_down_packed_kernel(const uint32_t* words, int32_t n,
                                                                        const uint32_t* sums, const uint32_t* valid,
                                                                        const uint4* erec, int ib, uint32_t* prefix,
                                                                        uint32_t* windows, int64_t nwin,
                                                                        uint32_t* total) {
    static_assert(kScanItems == 8, "two 16-byte loads and stores per lane");
    __shared__ uint32_t s_w[4];
    __shared__ uint64_t s_w64[4];
    const int64_t my = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    const int64_t nv = (int64_t)min(*valid, (uint32_t)n);
    const bool whole = my + kScanItems <= nv;   // the lane's 8 words in range: 16-byte accesses
    uint32_t wd[kScanItems];
    if (whole) {
        const uint4 lo = *reinterpret_cast<const uint4*>(words + my);
        const uint4 hi = *reinterpret_cast<const uint4*>(words + my + 4);
        wd[0] = lo.x; wd[1] = lo.y; wd[2] = lo.z; wd[3] = lo.w;
        wd[4] = hi.x; wd[5] = hi.y; wd[6] = hi.z; wd[7] = hi.w;
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) wd[k] = (my + k < nv) ? words[my + k] : 0u;
    }
    uint32_t v[kScanItems];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (my + k < nv) ? packed_count(wd[k], ib, erec) : 0u;
        sum += v[k];
    }
    uint32_t run = block_excl_256(sum, block_offset_256(sums, total, s_w64), s_w);
    uint32_t pf[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        pf[k] = run;
        run += v[k];
    }
    if (whole) {
        *reinterpret_cast<uint4*>(prefix + my) = make_uint4(pf[0], pf[1], pf[2], pf[3]);
        *reinterpret_cast<uint4*>(prefix + my + 4) = make_uint4(pf[4], pf[5], pf[6], pf[7]);
    }
    run = pf[0];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (my + k < nv) {
            if (!whole) prefix[my + k] = run;
            const uint64_t end = (uint64_t)run + v[k];
            for (uint64_t b = ((uint64_t)run + kEmitWindow - 1) / kEmitWindow; b * kEmitWindow < end && (int64_t)b < nwin;
                 ++b)
                windows[b] = (uint32_t)(my + k);   // nwin == 0: no table
        }
        run += v[k];
    }
}

int scan_scratch_words(int32_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

hipError_t launch_build_scan(uint32_t* data, int32_t n, uint32_t* total, uint32_t* scratch, hipStream_t s) {
    if (n <= 0) {
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, s, total, 0u, (int64_t)1);
        return hipGetLastError();
    }
    const int nb = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(scan_reduce_kernel, dim3(nb), dim3(kScanThreads), 0, s, data, n, scratch);
    hipLaunchKernelGGL(scan_down_kernel, dim3(nb), dim3(kScanThreads), 0, s, data, n, scratch, total);
    return hipGetLastError();
}

int packed_index_bits(int64_t n) {
    int b = 1;
    while (b < 31 && ((int64_t)1 << b) < n) ++b;
    return b;
}

hipError_t launch_build_scan_packed(const uint32_t* words, int32_t n, const uint32_t* valid, const uint4* erec,
                                    int32_t index_bits, uint32_t* total, uint32_t* scratch, uint32_t* prefix,
                                    uint32_t* windows, int64_t nwindows, hipStream_t s) {
    if (index_bits < 1 || index_bits > 31 || (n > 0 && ((int64_t)1 << index_bits) < n)) return hipErrorInvalidValue;
    if (n <= 0) {
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, s, total, 0u, (int64_t)1);
        return hipGetLastError();
    }
    if (!words || !valid || !erec || !scratch || !prefix || nwindows < 0 || (nwindows > 0 && !windows))
        return hipErrorInvalidValue;
    const int nb = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(scan_reduce_packed_kernel, dim3(nb), dim3(kScanThreads), 0, s, words, n, scratch, valid, erec,
                       index_bits);
    hipLaunchKernelGGL(scan_down_packed_kernel, dim3(nb), dim3(kScanThreads), 0, s, words, n, scratch, valid, erec,
                       index_bits, prefix, windows, nwindows, total);
    return hipGetLastError();
}

// The emit windows' first records from the scanned offsets (for a caller that sizes the table
// after the scan): windows[w] = the q whose entries [prefix[q], prefix[q + 1]) hold entry
// w * kEmitWindow (the last record's entries end at *total).
__global__ void emit_windows_kernel(const uint32_t* prefix, const uint32_t* valid, int32_t n, const uint32_t* total,
                                    uint32_t* windows, int64_t nwin) {
    const int64_t nv = (int64_t)min(*valid, (uint32_t)n);
    const uint32_t kt = *total;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nv; q += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t start = prefix[q];
        const uint64_t end = q + 1 < nv ? prefix[q + 1] : kt;
        for (uint64_t b = ((uint64_t)start + kEmitWindow - 1) / kEmitWindow; b * kEmitWindow < end && (int64_t)b < nwin;
             ++b)
            windows[b] = (uint32_t)q;
    }
}

hipError_t launch_emit_windows(const uint32_t* prefix, const uint32_t* valid, int32_t n, const uint32_t* total,
                               uint32_t* windows, int64_t nwindows, hipStream_t s) {
    if (n <= 0 || nwindows <= 0) return hipSuccess;
    if (!prefix || !valid || !total || !windows) return hipErrorInvalidValue;
    hipLaunchKernelGGL(emit_windows_kernel, dim3(grid_for(n)), dim3(256), 0, s, prefix, valid, n, total, windows,
                       nwindows);
    return hipGetLastError();
}

// The reference's (tile << 32 | depth word) keys and index values of a tile
// list held as (index << 32 | tile) entries: the depth word of an entry is
// that of its Gaussian in view 0.
__global__ void expand_keys_kernel(const uint64_t* entries, const float* clip_z, uint64_t* keys, int32_t* values,
                                   int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t e = entries[i];
        const int32_t g = (int32_t)(e >> 32);
        keys[i] = ((e & 0xffffffffull) << 32) | depth_word(clip_z[g]);
        values[i] = g;
    }
}

hipError_t launch_expand_keys(const uint64_t* entries, const float* clip_z, uint64_t* keys, int32_t* values,
                              int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(expand_keys_kernel, dim3(grid_for(n)), dim3(256), 0, s, entries, clip_z, keys, values, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Tile ranges: core/gaussian_render.cu:634-660; the -1 pre-fill is
// app/main.cpp:365.
// ---------------------------------------------------------------------------
__global__ void fill_i32_kernel(int32_t* p, int32_t value, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

hipError_t launch_fill_i32(int32_t* p, int32_t value, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, value, n);
    return hipGetLastError();
}

// Each lane takes 4 consecutive entries (two 16-byte loads) and the entry before them.
constexpr int kRangeItems = 4;

__global__ __launch_bounds__(256) void ranges_kernel(RangeArgs a) {
    if (a.size_dev && *a.size_dev > (uint32_t)a.size) return;   // the list did not fit: nothing to range
    const int64_t size = a.size_dev ? (int64_t)*a.size_dev : (int64_t)a.size;
    const int64_t i0 = kRangeItems * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (i0 >= size) return;
    // tile ids at or past `tiles` never occur in a well-formed list; they are ignored, not written
    const uint32_t nt = (uint32_t)a.tiles;
    const int sh = a.tile_shift;
    uint64_t e[kRangeItems];
    // 16-byte loads when the list is 16-byte aligned (the library's lists are; a caller's list of
    // the stage API need only be 8-byte aligned)
    if (i0 + kRangeItems <= size && ((uintptr_t)a.keys & 15u) == 0) {
        const ulonglong2* p = reinterpret_cast<const ulonglong2*>(a.keys + i0);
        const ulonglong2 u = p[0], v = p[1];
        e[0] = u.x;
        e[1] = u.y;
        e[2] = v.x;
        e[3] = v.y;
    } else {
#pragma unroll
        for (int j = 0; j < kRangeItems; ++j) e[j] = i0 + j < size ? a.keys[i0 + j] : 0ull;
    }
    uint32_t prev = i0 > 0 ? (uint32_t)(a.keys[i0 - 1] >> sh) : 0u;
#pragma unroll
    for (int j = 0; j < kRangeItems; ++j) {
        const int64_t i = i0 + j;
        if (i < size) {
            const uint32_t cur = (uint32_t)(e[j] >> sh);
            if (i == 0) {
                if (cur < nt) a.tile_range[(size_t)cur * 2] = 0;
            } else if (cur != prev) {
                if (prev < nt) a.tile_range[(size_t)prev * 2 + 1] = (int32_t)i;
                if (cur < nt) a.tile_range[(size_t)cur * 2] = (int32_t)i;
            }
            if (i == size - 1 && cur < nt) a.tile_range[(size_t)cur * 2 + 1] = (int32_t)size;
            prev = cur;
        }
    }
}

hipError_t launch_ranges(const RangeArgs& a, hipStream_t s) {
    if (a.size <= 0) return hipSuccess;
    if (!a.keys || (a.tile_shift != 0 && a.tile_shift != 32)) return hipErrorInvalidValue;
    const int64_t lanes = ((int64_t)a.size + kRangeItems - 1) / kRangeItems;
    hipLaunchKernelGGL(ranges_kernel, dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Raster: core/gaussian_render.cu:679-759.  Pixel (px, py) sits at clip
// (px * 2/W - 1, py * 2/H - 1) (no half-pixel offset, :701-704).  Entries are
// blended front to back; after every 32 entries of the tile the whole tile
// stops if no pixel has T > 0.02 (:739-748).  Pixels of edge tiles outside the
// frame take part in the vote but are not written.  The blend arithmetic is
// the one of gr::blend_exponent / blend_density below:
//   w = T * alpha;  C = fma(rgb, w, C);  T = T - w    (= C += rgb*T*alpha, T *= 1-alpha)
// alpha is 0 for a pair whose exponent is below -25 (or NaN).  While T is finite
// and the entry's colour is finite, such a pair leaves C and T as they are (up
// to the sign of a zero, which the output cannot show), so the lane simply sits
// it out: the blend runs under the mask of the lanes at or above the cutoff,
// and a wave whose lanes are all below skips the entry.  T stays in [0, 1] as
// long as every opacity is in [0, 1].  An entry whose colour is not finite (a
// NaN or infinite channel times a zero weight is NaN, which the reference's
// blend (:729-736) adds to every pixel it reaches) or whose opacity is outside
// [0, 1] (after which T need not be finite) is blended by every lane with the
// explicit zero, and so is every later entry of the tile.
// ---------------------------------------------------------------------------
constexpr int kRasterThreads = 256;
// Entries staged through LDS per round (a multiple of the 32-entry vote).  Most tiles stop after
// 100-400 entries, so smaller rounds gather fewer entries that are never blended: at 6M/1080p,
// 256 / 128 / 64 entries per round, raster 210 / 207 / 211 us.
#ifndef GR_RASTER_STAGE
#define GR_RASTER_STAGE 128
#endif
constexpr int kRasterStage = GR_RASTER_STAGE;
static_assert(kRasterStage % 32 == 0 && kRasterStage <= kRasterThreads, "stage whole vote batches, one per lane");
constexpr int kRasterBatch = 32;

// alpha = opacity * 2^min(e, 0) when e >= -25, else 0 (a pair whose density is below 2^-25
// contributes nothing; NaN contributes nothing).  The clamp of e to [-26, 0] is one v_med3_f32
// (min(e, 0) for the e >= -25 that reach it) and keeps floor(e) in int range.
__device__ __forceinline__ float blend_density(float e) {
    return exp2_poly(__builtin_amdgcn_fmed3f(e, -26.0f, 0.0f));
}

// One staged entry: the exponent's operands in the first 16 bytes and qc right after them, the
// blend's in the next 16.
struct __align__(16) RasterEntry {
    float4 a;   // cx, cy, qa, qb
    float4 b;   // opacity, r, g, b
    float qc;
    float pad[3];
};

__global__ __launch_bounds__(kRasterThreads) void raster_kernel(RasterArgs a) {
    __shared__ RasterEntry s_e[kRasterStage];
    __shared__ int s_alive[2][4];

    const int bid = blockIdx.x;
    // workgroup b shades the b-th tile of the order (heaviest first, from tile_order_kernel):
    // consecutive workgroups go to the 8 XCDs in turn, so the heavy tiles start first and spread
    // over the XCDs, and the light ones fill in at the end (index order measured 213 us, a
    // contiguous eighth of the tiles per XCD 219 us, heaviest first 208-210 us at 6M/1080p)
    const int tile = a.order ? (int)a.order[bid] : bid;
    if (tile >= a.tiles) return;   // uniform over the workgroup
    int view = 0;
    if (a.views) {
        while (view + 1 < a.nviews && a.views[view + 1].tile_base <= tile) ++view;
    }
    const ViewDesc& vd = view_of(view, a.v0, a.views);
    const FrameGeom& g = vd.geom;
    const int local = tile - vd.tile_base;
    const int tx = local % g.tiles_x, ty = local / g.tiles_x;
    const size_t vo = (size_t)view * a.count;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int px = tx * kTileSize + (w & 1) * 8 + (lane & 7);
    const int py = ty * kTileSize + (w >> 1) * 8 + (lane >> 3);
    const float pcx = (float)px * g.texel_clip_x - 1.0f;
    const float pcy = (float)py * g.texel_clip_y - 1.0f;

    const int start = a.tile_range[2 * tile];
    const int end = a.tile_range[2 * tile + 1];
    const int n = end - start;

    float C0 = 0.0f, C1 = 0.0f, C2 = 0.0f, T = 1.0f;
    int parity = 0;
    if (n > 0) {
        bool stop = false;
        bool full = false;   // from the round with the tile's first flagged entry on: the full rule
        for (int r0 = 0; r0 < n && !stop; r0 += kRasterStage) {
            const int nr = min(kRasterStage, n - r0);
            __syncthreads();
            bool flagged = false;
            if ((int)threadIdx.x < nr) {
                const int k = start + r0 + threadIdx.x;
                const size_t gi = vo + (size_t)(a.values ? a.values[k] : (int32_t)(a.entries[k] >> 32));
                const float2 c = a.clip_xy[gi];
                const float4 q = a.conic[gi];
                const float4 col = a.color[gi];
                flagged = !(isfinite(col.x) && isfinite(col.y) && isfinite(col.z) && col.w >= 0.0f &&
                            col.w <= 1.0f);
                s_e[threadIdx.x].a = make_float4(c.x, c.y, q.x * kHalfLog2e, q.y * kLog2e);
                s_e[threadIdx.x].b = make_float4(col.w, col.x, col.y, col.z);
                s_e[threadIdx.x].qc = q.z * kHalfLog2e;
            }
            // the round's entries before the flagged one may take the full rule too: for them
            // both rules give the same C and T
            if (__syncthreads_or(flagged ? 1 : 0)) full = true;
            for (int b0 = 0; b0 < nr; b0 += kRasterBatch) {
                const int b1 = min(b0 + kRasterBatch, nr);
                auto exponent = [&](const RasterEntry& en) {
                    const float4 ea = en.a;
                    const float dx = pcx - ea.x, dy = pcy - ea.y;
                    return blend_exponent(ea.z, ea.w, en.qc, dx, dy);
                };
                if (full) {
                    // every lane applies the rule, alpha 0 below the cutoff included
                    for (int j = b0; j < b1; ++j) {
                        const float e = exponent(s_e[j]);
                        const float4 eb = s_e[j].b;
                        const float wgt = T * ((e >= kMinExponent) ? eb.x * blend_density(e) : 0.0f);
                        C0 = fmaf(eb.y, wgt, C0);
                        C1 = fmaf(eb.z, wgt, C1);
                        C2 = fmaf(eb.w, wgt, C2);
                        T = T - wgt;
                    }
                } else {
                    auto blend = [&](const RasterEntry& en) {
                        const float e = exponent(en);
                        if (e >= kMinExponent) {   // lanes below the cutoff (or NaN) sit the entry out
                            const float4 eb = en.b;
                            const float wgt = T * (eb.x * blend_density(e));
                            C0 = fmaf(eb.y, wgt, C0);
                            C1 = fmaf(eb.z, wgt, C1);
                            C2 = fmaf(eb.w, wgt, C2);
                            T = T - wgt;
                        }
                    };
                    int j = b0;
                    for (; j + 1 < b1; j += 2) {
                        // the pair's LDS address in a vector register (an empty asm keeps the
                        // compiler from rebuilding it from the scalar index before every read)
                        int jv = j * (int)sizeof(RasterEntry);
                        asm volatile("" : "+v"(jv));
                        const RasterEntry* pe =
                            reinterpret_cast<const RasterEntry*>(reinterpret_cast<const char*>(s_e) + jv);
                        blend(pe[0]);
                        blend(pe[1]);
                    }
                    if (j < b1) blend(s_e[j]);
                }
                const bool alive = __ballot(T > kTransmittanceStop) != 0ull;
                if (lane == 0) s_alive[parity][w] = alive ? 1 : 0;
                __syncthreads();
                const int any = s_alive[parity][0] | s_alive[parity][1] | s_alive[parity][2] |
                                s_alive[parity][3];
                parity ^= 1;
                if (!any) { stop = true; break; }
            }
        }
    }
    if (px < g.width && py < g.height) {
        uchar4 o;
        if (n > 0) {
            o.x = (unsigned char)f2u_sat(gclamp(C0, 0.0f, 1.0f) * 255.0f);
            o.y = (unsigned char)f2u_sat(gclamp(C1, 0.0f, 1.0f) * 255.0f);
            o.z = (unsigned char)f2u_sat(gclamp(C2, 0.0f, 1.0f) * 255.0f);
            o.w = 255;
        } else {
            o = make_uchar4(0, 0, 0, 0);
        }
        vd.out[(size_t)py * g.width + px] = o;
    }
}

// Raster order: the tiles by decreasing entry count, in classes of powers of two (order within a
// class unspecified), so that the heaviest tiles start first and the lightest fill in at the end.
// One workgroup; tiles without entries (range -1) come last.
// Each lane loads the ranges of 8 tiles before it counts or places them, so that a pass waits
// for one round of loads per 8 192 tiles instead of one per 1 024 (a batch of views has up to
// 2^16 tiles).  (One lane per class and wave doing the LDS atomic, the wave's class peers found
// with 6 ballots, measured 12.1 against 8.3 us at 8 160 tiles: the one workgroup is bound by its
// instructions and load latency, not the atomics.)
constexpr int kOrderUnroll = 8;
__global__ __launch_bounds__(1024) void tile_order_kernel(const int32_t* range, int32_t tiles, uint32_t* order) {
    __shared__ uint32_t s_n[33];
    __shared__ uint32_t s_at[33];
    if (threadIdx.x < 33) s_n[threadIdx.x] = 0u;
    __syncthreads();
    const int2* rg = reinterpret_cast<const int2*>(range);
    auto classes = [&](int t0, int* c) {   // leading zeros of the entry count: 0 = most, 32 = none
        if (t0 + (kOrderUnroll - 1) * (int)blockDim.x < tiles) {   // all 8 in range: loads issued together
            int2 r[kOrderUnroll];
#pragma unroll
            for (int k = 0; k < kOrderUnroll; ++k) r[k] = rg[t0 + k * (int)blockDim.x];
#pragma unroll
            for (int k = 0; k < kOrderUnroll; ++k) c[k] = __clz(max(r[k].y - r[k].x, 0));
        } else {
#pragma unroll
            for (int k = 0; k < kOrderUnroll; ++k) {
                const int t = t0 + k * (int)blockDim.x;
                c[k] = t < tiles ? __clz(max(rg[t].y - rg[t].x, 0)) : -1;
            }
        }
    };
    const int step = kOrderUnroll * (int)blockDim.x;
    for (int t0 = threadIdx.x; t0 < tiles; t0 += step) {
        int c[kOrderUnroll];
        classes(t0, c);
#pragma unroll
        for (int k = 0; k < kOrderUnroll; ++k)
            if (c[k] >= 0) atomicAdd(&s_n[c[k]], 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0u;
        for (int c = 0; c < 33; ++c) {
            s_at[c] = run;
            run += s_n[c];
        }
    }
    __syncthreads();
    for (int t0 = threadIdx.x; t0 < tiles; t0 += step) {
        int c[kOrderUnroll];
        classes(t0, c);
#pragma unroll
        for (int k = 0; k < kOrderUnroll; ++k)
            if (c[k] >= 0) order[atomicAdd(&s_at[c[k]], 1u)] = (uint32_t)(t0 + k * (int)blockDim.x);
    }
}

hipError_t launch_tile_order(const int32_t* tile_range, int32_t tiles, uint32_t* order, hipStream_t s) {
    if (tiles <= 0) return hipSuccess;
    if (!tile_range || !order) return hipErrorInvalidValue;
    hipLaunchKernelGGL(tile_order_kernel, dim3(1), dim3(1024), 0, s, tile_range, tiles, order);
    return hipGetLastError();
}

hipError_t launch_raster(const RasterArgs& a, hipStream_t s) {
    if (a.tiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(raster_kernel, dim3((unsigned)a.tiles), dim3(kRasterThreads), 0, s, a);
    return hipGetLastError();
}

}  // namespace gr
