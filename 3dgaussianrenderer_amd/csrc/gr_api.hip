// gr_api.hip -- host side of the C-ABI (include/gr_capi.h): contexts, timers,
// the stage launchers of core/gaussian_render.cuh:113-126 and the object-level
// renderer of core/renderer.{h,cpp}.  No entry point exits the process; every
// failure returns a gr_status and leaves a message for gr_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gr_capi.h"
#include "gr_internal.hpp"
#include "gr_kernels.hpp"

using namespace gr;

namespace {

thread_local std::string g_last_error;

gr_status fail(gr_status st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return st;
}

#define GR_HIP(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(e_ == hipErrorOutOfMemory ? GR_ERR_OUT_OF_MEMORY : GR_ERR_HIP,      \
                        "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

#define GR_CHECK(st)                      \
    do {                                  \
        gr_status s_ = (st);              \
        if (s_ != GR_OK) return s_;       \
    } while (0)

template <typename T>
gr_status grow(T*& p, size_t& cap, size_t need, double headroom = 1.0) {
    if (need <= cap && p != nullptr) return GR_OK;
    size_t n = (size_t)((double)need * headroom);
    if (n < need) n = need;
    if (n == 0) n = 1;
    if (p) GR_HIP(hipFree(p));
    p = nullptr;
    cap = 0;
    GR_HIP(hipMalloc((void**)&p, n * sizeof(T)));
    cap = n;
    return GR_OK;
}

}  // namespace

namespace gr_detail {
void set_last_error(const std::string& msg) { g_last_error = msg; }
std::string get_last_error() { return g_last_error; }
}  // namespace gr_detail

static gr_status timer_start(gr_context* ctx, gr_timer* t) {
    if (!t) return GR_OK;
    GR_HIP(hipEventRecord(t->start, ctx->stream));
    return GR_OK;
}
static gr_status timer_stop(gr_context* ctx, gr_timer* t) {
    if (!t) return GR_OK;
    GR_HIP(hipEventRecord(t->stop, ctx->stream));
    t->recorded = true;
    return GR_OK;
}

static FrameGeom ctx_geom(const gr_context* ctx) {
    return make_geom(ctx->args.width, ctx->args.height);
}

// The one view of the stage-level API (tile ids start at 0).
static ViewDesc ctx_view(const gr_context* ctx) {
    ViewDesc v{};
    v.cam = camera_args(ctx->args.camera_data);
    v.geom = ctx_geom(ctx);
    v.tile_base = 0;
    v.tiles = v.geom.tiles_x * v.geom.tiles_y;
    v.out = reinterpret_cast<uchar4*>(ctx->args.back_buffer);
    return v;
}

extern "C" {

const char* gr_last_error(void) { return g_last_error.c_str(); }
const char* gr_version(void) { return "gr-mi355x 0.1 (gfx950)"; }

gr_status gr_context_create(gr_context** out, void* stream) { return gr_context_create_ex(out, stream, 0); }

gr_status gr_context_create_ex(gr_context** out, void* stream, int32_t flags) {
    if (!out) return fail(GR_ERR_INVALID_ARGUMENT, "gr_context_create: out is NULL");
    auto* c = new gr_context();
    if (stream || (flags & GR_CONTEXT_BORROW_STREAM)) {
        c->stream = static_cast<hipStream_t>(stream);
    } else {
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            return fail(GR_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
        }
        c->own_stream = true;
    }
    hipError_t e = hipHostMalloc((void**)&c->pinned, 64, hipHostMallocDefault);
    if (e != hipSuccess) {
        if (c->own_stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(GR_ERR_HIP, "hipHostMalloc failed: %s", hipGetErrorString(e));
    }
    *out = c;
    return GR_OK;
}

gr_status gr_context_destroy(gr_context* ctx) {
    if (!ctx) return GR_OK;
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->batch_counts) (void)hipFree(ctx->batch_counts);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return GR_OK;
}

gr_status gr_context_synchronize(gr_context* ctx) {
    if (!ctx) return fail(GR_ERR_INVALID_ARGUMENT, "context is NULL");
    GR_HIP(hipStreamSynchronize(ctx->stream));
    return GR_OK;
}

gr_status gr_timer_create(gr_timer** out) {
    if (!out) return fail(GR_ERR_INVALID_ARGUMENT, "gr_timer_create: out is NULL");
    auto* t = new gr_timer();
    // timing only (the launchers' results reach the host through the caller's own copies and
    // synchronisation), so the events skip the system-scope cache fence
    if (hipEventCreateWithFlags(&t->start, hipEventDisableSystemFence) != hipSuccess ||
        hipEventCreateWithFlags(&t->stop, hipEventDisableSystemFence) != hipSuccess) {
        delete t;
        return fail(GR_ERR_HIP, "hipEventCreate failed");
    }
    *out = t;
    return GR_OK;
}

gr_status gr_timer_destroy(gr_timer* t) {
    if (!t) return GR_OK;
    (void)hipEventDestroy(t->start);
    (void)hipEventDestroy(t->stop);
    delete t;
    return GR_OK;
}

gr_status gr_timer_elapsed_ms(gr_timer* t, float* out_ms) {
    if (!t || !out_ms) return fail(GR_ERR_INVALID_ARGUMENT, "gr_timer_elapsed_ms: NULL argument");
    if (!t->recorded) return fail(GR_ERR_INVALID_ARGUMENT, "timer was never stopped");
    GR_HIP(hipEventSynchronize(t->stop));
    GR_HIP(hipEventElapsedTime(out_ms, t->start, t->stop));
    return GR_OK;
}

gr_status gr_set_global_args(gr_context* ctx, const gr_global_args* args) {
    if (!ctx || !args) return fail(GR_ERR_INVALID_ARGUMENT, "gr_set_global_args: NULL argument");
    if (args->width <= 0 || args->height <= 0)
        return fail(GR_ERR_INVALID_ARGUMENT, "frame size %dx%d is invalid", args->width, args->height);
    if (args->spherical_harmonics_degree < 0 || args->spherical_harmonics_degree > 3)
        return fail(GR_ERR_INVALID_ARGUMENT, "SH degree %d is outside 0..3",
                    args->spherical_harmonics_degree);
    const int want = 3 * (args->spherical_harmonics_degree + 1) * (args->spherical_harmonics_degree + 1);
    if (args->sh_precision < GR_PRECISION_FP32 || args->sh_precision > GR_PRECISION_BF16)
        return fail(GR_ERR_INVALID_ARGUMENT, "unknown precision %d", args->sh_precision);
    if (args->spherical_harmonics_count != want)
        return fail(GR_ERR_INVALID_ARGUMENT, "spherical_harmonics_count %d != 3(d+1)^2 = %d",
                    args->spherical_harmonics_count, want);
    ctx->args = *args;
    ctx->has_args = true;
    return GR_OK;
}

gr_status gr_set_tile_list_args(gr_context* ctx, const gr_tile_list_args* args) {
    if (!ctx || !args) return fail(GR_ERR_INVALID_ARGUMENT, "gr_set_tile_list_args: NULL argument");
    ctx->tl = *args;
    ctx->has_tl = true;
    return GR_OK;
}

gr_status gr_evaluate_spherical_harmonics(gr_context* ctx, gr_timer* timer, int32_t count) {
    if (!ctx || !ctx->has_args) return fail(GR_ERR_INVALID_ARGUMENT, "global args not set");
    const gr_global_args& g = ctx->args;
    if (count < 0 || count > g.splat_count) return fail(GR_ERR_INVALID_ARGUMENT, "bad count %d", count);
    if (count > 0 && (!g.position || !g.color || !g.spherical_harmonics))
        return fail(GR_ERR_INVALID_ARGUMENT, "SH stage needs position, color and SH buffers");
    if (g.sh_precision != GR_PRECISION_FP32 && count > 0 && !g.packed_scale_opacity)
        return fail(GR_ERR_INVALID_ARGUMENT, "precision form needs packed_scale_opacity");
    ShArgs a;
    a.position = reinterpret_cast<const float4*>(g.position);
    a.sh = g.spherical_harmonics;
    a.packed = reinterpret_cast<const uint2*>(g.packed_scale_opacity);
    a.color = reinterpret_cast<float4*>(g.color);
    a.count = count;
    a.sh_count = g.spherical_harmonics_count;
    a.nviews = 1;
    a.v0 = ctx_view(ctx);
    a.views = nullptr;
    GR_CHECK(timer_start(ctx, timer));
    GR_HIP(launch_sh(a, g.spherical_harmonics_degree, g.sh_precision, ctx->stream));
    GR_CHECK(timer_stop(ctx, timer));
    return GR_OK;
}

gr_status gr_evaluate_splat_clip_data(gr_context* ctx, gr_timer* timer, int32_t count) {
    if (!ctx || !ctx->has_args) return fail(GR_ERR_INVALID_ARGUMENT, "global args not set");
    const gr_global_args& g = ctx->args;
    if (count < 0 || count > g.splat_count) return fail(GR_ERR_INVALID_ARGUMENT, "bad count %d", count);
    const bool packed = g.sh_precision != GR_PRECISION_FP32;
    if (count > 0 && (!g.position || (!packed && !g.scale_and_rotation) || (packed && !g.packed_scale_opacity) ||
                      !g.position_clip_xy || !g.position_clip_z || !g.screen_ellipse || !g.conic))
        return fail(GR_ERR_INVALID_ARGUMENT, "projection stage is missing a buffer");
    ProjectArgs a{};
    a.position = reinterpret_cast<const float4*>(g.position);
    a.scale_rotation = reinterpret_cast<const float4*>(g.scale_and_rotation);
    a.packed = reinterpret_cast<const uint2*>(g.packed_scale_opacity);
    a.precision = g.sh_precision;
    a.clip_xy = reinterpret_cast<float2*>(g.position_clip_xy);
    a.clip_z = g.position_clip_z;
    a.ellipse = reinterpret_cast<float4*>(g.screen_ellipse);
    a.conic = reinterpret_cast<float4*>(g.conic);
    a.count = count;
    a.nviews = 1;
    a.v0 = ctx_view(ctx);
    a.views = nullptr;
    GR_CHECK(timer_start(ctx, timer));
    GR_HIP(launch_project(a, ctx->stream));
    GR_CHECK(timer_stop(ctx, timer));
    return GR_OK;
}

gr_status gr_build_tile_list(gr_context* ctx, gr_timer* timer, int32_t num_blocks,
                             int32_t tile_list_capacity, int32_t* out_size, int64_t* out_required) {
    (void)num_blocks;
    if (!ctx || !ctx->has_args || !ctx->has_tl)
        return fail(GR_ERR_INVALID_ARGUMENT, "global/tile-list args not set");
    if (tile_list_capacity < 0) return fail(GR_ERR_INVALID_ARGUMENT, "negative capacity");
    const gr_global_args& g = ctx->args;
    const int n = g.splat_count;
    const int batches = (n + kBuildBatch - 1) / kBuildBatch;
    GR_CHECK(grow(ctx->batch_counts, ctx->batch_cap, (size_t)batches + 1 + scan_scratch_words(batches)));
    BuildArgs a{};
    a.clip_xy = reinterpret_cast<const float2*>(g.position_clip_xy);
    a.clip_z = g.position_clip_z;
    a.ellipse = reinterpret_cast<const float4*>(g.screen_ellipse);
    a.batch_counts = ctx->batch_counts;
    a.keys = ctx->tl.keys;
    a.values = ctx->tl.values;
    a.count = n;
    a.capacity = tile_list_capacity;
    a.batches = batches;
    a.nviews = 1;
    a.v0 = ctx_view(ctx);
    a.views = nullptr;
    if (a.v0.tiles >= (1 << 20))
        return fail(GR_ERR_INVALID_ARGUMENT, "frame has too many tiles");
    GR_CHECK(timer_start(ctx, timer));
    uint32_t total = 0;
    if (n > 0) {
        GR_HIP(launch_build_count(a, ctx->stream));
        GR_HIP(launch_build_scan(ctx->batch_counts, batches, ctx->batch_counts + batches,
                                 ctx->batch_counts + batches + 1, ctx->stream));
        GR_HIP(hipMemcpyAsync(ctx->pinned, ctx->batch_counts + batches, sizeof(uint32_t),
                              hipMemcpyDeviceToHost, ctx->stream));
        GR_HIP(hipStreamSynchronize(ctx->stream));
        total = ctx->pinned[0];
        if (total > 0 && tile_list_capacity > 0) {
            if (!a.keys || !a.values) return fail(GR_ERR_INVALID_ARGUMENT, "tile-list buffers are NULL");
            GR_HIP(launch_build_emit(a, ctx->stream));
        }
    }
    GR_CHECK(timer_stop(ctx, timer));
    const int64_t k = (int64_t)total;
    if (out_required) *out_required = k;
    if (out_size) *out_size = (int32_t)std::min<int64_t>(k, tile_list_capacity);
    return GR_OK;
}

gr_status gr_device_free(void* p) {
    if (p) GR_HIP(hipFree(p));
    return GR_OK;
}

gr_status gr_radix_sort_pairs(gr_context* ctx, int32_t count, int32_t begin_bit, int32_t end_bit,
                              void** temp, size_t* temp_bytes, gr_double_buffer_u64* keys,
                              gr_double_buffer_i32* values) {
    if (!ctx || !temp || !temp_bytes || !keys || !values)
        return fail(GR_ERR_INVALID_ARGUMENT, "gr_radix_sort_pairs: NULL argument");
    if (count < 0 || count >= (1 << 30)) return fail(GR_ERR_INVALID_ARGUMENT, "count %d out of range", count);
    if (begin_bit < 0 || end_bit > 64 || end_bit - begin_bit > 64)
        return fail(GR_ERR_INVALID_ARGUMENT, "bad bit range [%d, %d)", begin_bit, end_bit);
    const size_t need = radix_sort_temp_bytes(count, begin_bit, end_bit);
    if (need > *temp_bytes || *temp == nullptr) {
        if (*temp) GR_HIP(hipFree(*temp));
        *temp = nullptr;
        *temp_bytes = 0;
        GR_HIP(hipMalloc(temp, need));
        *temp_bytes = need;
    }
    uint64_t* kb[2] = {keys->buffers[0], keys->buffers[1]};
    int32_t* vb[2] = {values->buffers[0], values->buffers[1]};
    int32_t sel = keys->selector;
    if (values->selector != sel) return fail(GR_ERR_INVALID_ARGUMENT, "key/value selectors differ");
    GR_HIP(radix_sort_pairs(kb, vb, &sel, count, nullptr, begin_bit, end_bit, *temp, *temp_bytes, ctx->stream));
    keys->selector = sel;
    values->selector = sel;
    return GR_OK;
}

gr_status gr_sort_tile_list(gr_context* ctx, gr_timer* timer, int32_t size, void** temp,
                            size_t* temp_bytes, gr_double_buffer_u64* keys, gr_double_buffer_i32* values) {
    if (!ctx || !ctx->has_args) return fail(GR_ERR_INVALID_ARGUMENT, "global args not set");
    const FrameGeom g = ctx_geom(ctx);
    const int end_bit = 32 + tile_bits_for(g.tiles_x * g.tiles_y);
    GR_CHECK(timer_start(ctx, timer));
    GR_CHECK(gr_radix_sort_pairs(ctx, size, 0, end_bit, temp, temp_bytes, keys, values));
    GR_CHECK(timer_stop(ctx, timer));
    return GR_OK;
}

}  // extern "C"

// ---- sortTileList without allocation ------------------------------------
// The scratch of gr_sort_tile_list_static: digit histograms, partition tickets and the two
// look-back status regions of the onesweep sort, in one file-scope device array (every GPU the
// library runs on holds its own copy until the process ends).  Host code never passes the
// array's name to a kernel -- on the host it is only a stand-in -- but the device address that
// hipGetSymbolAddress returns for the current device, looked up once per device.
constexpr size_t kStaticSortScratchBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticSortScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first call per device).
gr_status static_scratch(void** out) {
    int dev = 0;
    GR_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(GR_ERR_INVALID_ARGUMENT, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        GR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(GR_ERR_HIP, "hipGetSymbolAddress returned NULL for the sort scratch");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return GR_OK;
}

// Largest pair count whose sort scratch fits the static array (the layout grows with the count).
int32_t static_sort_max_pairs() {
    int64_t lo = 1, hi = (int64_t)1 << 30;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) / 2;
        if (radix_sort_temp_bytes((int32_t)std::min<int64_t>(mid, (1 << 30) - 1), 0, 64) <= kStaticSortScratchBytes)
            lo = mid;
        else
            hi = mid - 1;
    }
    return (int32_t)std::min<int64_t>(lo, (1 << 30) - 1);
}
}  // namespace

extern "C" {

gr_status gr_sort_tile_list_static(gr_context* ctx, gr_timer* timer, int32_t size, gr_double_buffer_u64* keys,
                                   gr_double_buffer_i32* values, int32_t* max_pairs) {
    static const int32_t kMax = static_sort_max_pairs();
    if (max_pairs) *max_pairs = kMax;
    if (!ctx || !ctx->has_args) return fail(GR_ERR_INVALID_ARGUMENT, "global args not set");
    if (!keys || !values) return fail(GR_ERR_INVALID_ARGUMENT, "gr_sort_tile_list_static: NULL argument");
    if (size < 0) return fail(GR_ERR_INVALID_ARGUMENT, "negative tile-list size %d", size);
    if (size > kMax)
        return fail(GR_ERR_CAPACITY,
                    "tile list of %d pairs exceeds the %d pairs the static sort scratch (%zu bytes) holds",
                    size, kMax, kStaticSortScratchBytes);
    if (values->selector != keys->selector) return fail(GR_ERR_INVALID_ARGUMENT, "key/value selectors differ");
    void* temp = nullptr;
    GR_CHECK(static_scratch(&temp));
    const FrameGeom g = ctx_geom(ctx);
    const int end_bit = 32 + tile_bits_for(g.tiles_x * g.tiles_y);
    uint64_t* kb[2] = {keys->buffers[0], keys->buffers[1]};
    int32_t* vb[2] = {values->buffers[0], values->buffers[1]};
    int32_t sel = keys->selector;
    GR_CHECK(timer_start(ctx, timer));
    GR_HIP(radix_sort_pairs(kb, vb, &sel, size, nullptr, 0, end_bit, temp, kStaticSortScratchBytes, ctx->stream));
    GR_CHECK(timer_stop(ctx, timer));
    keys->selector = sel;
    values->selector = sel;
    return GR_OK;
}

gr_status gr_evaluate_tile_range(gr_context* ctx, gr_timer* timer, int32_t size) {
    if (!ctx || !ctx->has_args || !ctx->has_tl)
        return fail(GR_ERR_INVALID_ARGUMENT, "global/tile-list args not set");
    const FrameGeom g = ctx_geom(ctx);
    const int tiles = g.tiles_x * g.tiles_y;
    if (!ctx->args.tile_range) return fail(GR_ERR_INVALID_ARGUMENT, "tile_range buffer is NULL");
    GR_CHECK(timer_start(ctx, timer));
    GR_HIP(launch_fill_i32(ctx->args.tile_range, -1, 2 * (int64_t)tiles, ctx->stream));
    if (size > 0) {
        RangeArgs a{};
        a.keys = ctx->tl.keys;
        a.tile_shift = 32;
        a.tile_range = ctx->args.tile_range;
        a.size_dev = nullptr;
        a.size = size;
        a.tiles = tiles;
        GR_HIP(launch_ranges(a, ctx->stream));
    }
    GR_CHECK(timer_stop(ctx, timer));
    return GR_OK;
}

gr_status gr_rasterize_tile(gr_context* ctx, gr_timer* timer) {
    if (!ctx || !ctx->has_args || !ctx->has_tl)
        return fail(GR_ERR_INVALID_ARGUMENT, "global/tile-list args not set");
    const gr_global_args& g = ctx->args;
    if (!g.back_buffer || !g.tile_range) return fail(GR_ERR_INVALID_ARGUMENT, "raster buffers are NULL");
    // the kernel gathers every entry's Gaussian through values[] and reads its clip position,
    // conic and colour: a missing one is refused here, before anything is launched
    if (!ctx->tl.values || !g.position_clip_xy || !g.conic || !g.color)
        return fail(GR_ERR_INVALID_ARGUMENT, "raster needs tile-list values, clip positions, conics and colours");
    RasterArgs a{};
    a.v0 = ctx_view(ctx);
    a.views = nullptr;
    a.nviews = 1;
    a.count = g.splat_count;
    a.tiles = a.v0.tiles;
    a.tile_range = g.tile_range;
    a.values = ctx->tl.values;
    a.clip_xy = reinterpret_cast<const float2*>(g.position_clip_xy);
    a.conic = reinterpret_cast<const float4*>(g.conic);
    a.color = reinterpret_cast<const float4*>(g.color);
    GR_CHECK(timer_start(ctx, timer));
    GR_HIP(launch_raster(a, ctx->stream));
    GR_CHECK(timer_stop(ctx, timer));
    return GR_OK;
}

// ---- host helpers --------------------------------------------------------

gr_status gr_realign_spherical_harmonics(const float* src, float* dst, int32_t group_size,
                                         int32_t sh_count, int32_t splat_count) {
    if (!src || !dst || group_size <= 0 || sh_count <= 0 || splat_count < 0 || sh_count % 3 != 0)
        return fail(GR_ERR_INVALID_ARGUMENT, "gr_realign_spherical_harmonics: bad argument");
    const int groups = (splat_count + group_size - 1) / group_size;
    for (int grp = 0; grp < groups; ++grp) {
        const size_t start = (size_t)grp * group_size * sh_count;
        const int here = std::min(group_size, splat_count - grp * group_size);
        for (int i = 0; i < sh_count; ++i) {
            float* d = dst + start + (size_t)group_size * i;
            for (int k = 0; k < here; ++k) d[k] = src[start + (size_t)sh_count * k + i];
            for (int k = here; k < group_size; ++k) d[k] = 0.0f;
        }
    }
    return GR_OK;
}

uint16_t gr_float_to_half(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    const uint32_t sign = (u >> 16) & 0x8000u;
    const uint32_t a = u & 0x7fffffffu;
    if (a >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (a > 0x7f800000u ? 0x0200u : 0u));
    if (a >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // rounds to infinity
    if (a < 0x38800000u) {                                       // FP16 subnormal or zero
        float ax;
        std::memcpy(&ax, &a, 4);
        const float m = std::nearbyint(ax * 16777216.0f);         // units of 2^-24, ties to even
        return (uint16_t)(sign | (uint32_t)m);
    }
    uint32_t h = (((a >> 23) - 112u) << 10) | ((a >> 13) & 0x3ffu);
    const uint32_t rem = a & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
    return (uint16_t)(sign | h);
}

uint16_t gr_float_to_bf16(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

gr_status gr_pack_precision_model(const gr_model_host* m, int32_t precision, float* position_out,
                                  uint16_t* packed_out, uint16_t* sh_out) {
    if (!m || !position_out || !packed_out || m->num_gaussians < 0 || m->sh_degree < 0 || m->sh_degree > 3 ||
        (precision != GR_PRECISION_FP16 && precision != GR_PRECISION_BF16) ||
        (m->num_gaussians > 0 && (!m->position || !m->scale_and_rotation)) ||
        (m->sh_degree > 0 && m->num_gaussians > 0 && (!m->sh || !sh_out)))
        return fail(GR_ERR_INVALID_ARGUMENT, "gr_pack_precision_model: bad argument");
    auto cv = precision == GR_PRECISION_FP16 ? gr_float_to_half : gr_float_to_bf16;
    const int n = m->num_gaussians;
    for (int i = 0; i < n; ++i) {
        position_out[4 * i + 0] = m->position[4 * i + 0];
        position_out[4 * i + 1] = m->position[4 * i + 1];
        position_out[4 * i + 2] = m->position[4 * i + 2];
        position_out[4 * i + 3] = m->scale_and_rotation[4 * i + 3];  // packed quaternion bits
        packed_out[4 * i + 0] = cv(m->scale_and_rotation[4 * i + 0]);
        packed_out[4 * i + 1] = cv(m->scale_and_rotation[4 * i + 1]);
        packed_out[4 * i + 2] = cv(m->scale_and_rotation[4 * i + 2]);
        packed_out[4 * i + 3] = cv(m->position[4 * i + 3]);           // opacity
    }
    if (m->sh_degree > 0) {
        const int shc = 3 * (m->sh_degree + 1) * (m->sh_degree + 1);
        const int groups = (n + kShGroup - 1) / kShGroup;
        for (int grp = 0; grp < groups; ++grp) {
            const int here = std::min(kShGroup, n - grp * kShGroup);
            for (int j = 0; j < shc; ++j) {
                uint16_t* d = sh_out + ((size_t)grp * shc + j) * kShGroup;
                for (int k = 0; k < kShGroup; ++k)
                    d[k] = k < here ? cv(m->sh[((size_t)grp * kShGroup + k) * shc + j]) : (uint16_t)0;
            }
        }
    }
    return GR_OK;
}

uint32_t gr_encode_vec4(float x, float y, float z, float w) {
    const float v[4] = {x, y, z, w};
    uint32_t r = 0;
    for (int i = 0; i < 4; ++i) {
        const float c = gclamp(v[i], 0.0f, 1.0f);
        r |= (f2u_sat(c * 255.0f) & 0xffu) << (24 - 8 * i);
    }
    return r;
}

gr_status gr_tile_grid(int32_t width, int32_t height, int32_t* tiles_x, int32_t* tiles_y,
                       int32_t* sort_end_bit) {
    if (width <= 0 || height <= 0) return fail(GR_ERR_INVALID_ARGUMENT, "bad frame size");
    const FrameGeom g = make_geom(width, height);
    if (tiles_x) *tiles_x = g.tiles_x;
    if (tiles_y) *tiles_y = g.tiles_y;
    if (sort_end_bit) *sort_end_bit = 32 + tile_bits_for(g.tiles_x * g.tiles_y);
    return GR_OK;
}

// Camera: core/camera.cpp:7-55 (glm::lookAt = lookAtRH, ext/matrix_transform.inl:99-119;
// glm::perspective = perspectiveRH_NO, ext/matrix_clip_space.inl:249-262).
gr_status gr_camera_make(const float eye[3], const float center[3], float fov_y, int32_t width,
                         int32_t height, float z_near, float z_far, gr_camera_data* out) {
    if (!eye || !center || !out || width <= 0 || height <= 0 || !(z_far > z_near))
        return fail(GR_ERR_INVALID_ARGUMENT, "gr_camera_make: bad argument");
    const v3 e{eye[0], eye[1], eye[2]}, c{center[0], center[1], center[2]}, up{0.0f, 1.0f, 0.0f};
    const v3 f = normalize3(v3{c.x - e.x, c.y - e.y, c.z - e.z});
    const v3 s = normalize3(cross3(f, up));
    const v3 u = cross3(s, f);
    float V[16] = {0};
    V[0] = s.x; V[4] = s.y; V[8] = s.z;
    V[1] = u.x; V[5] = u.y; V[9] = u.z;
    V[2] = -f.x; V[6] = -f.y; V[10] = -f.z;
    V[12] = -dot3(s, e);
    V[13] = -dot3(u, e);
    V[14] = dot3(f, e);
    V[15] = 1.0f;
    const float aspect = (float)width / (float)height;
    const float tan_half = std::tan(fov_y / 2.0f);
    float P[16] = {0};
    P[0] = 1.0f / (aspect * tan_half);
    P[5] = 1.0f / tan_half;
    P[10] = -(z_far + z_near) / (z_far - z_near);
    P[11] = -1.0f;
    P[14] = -(2.0f * z_far * z_near) / (z_far - z_near);
    // view_projection = P * V (glm mat4 product: column c = P * V[c])
    float VP[16];
    const m4 pm = load_m4(P);
    for (int col = 0; col < 4; ++col) {
        const v4 r = mul44v(pm, v4{V[4 * col], V[4 * col + 1], V[4 * col + 2], V[4 * col + 3]});
        VP[4 * col] = r.x; VP[4 * col + 1] = r.y; VP[4 * col + 2] = r.z; VP[4 * col + 3] = r.w;
    }
    std::memcpy(out->view, V, sizeof(V));
    std::memcpy(out->projection, P, sizeof(P));
    std::memcpy(out->view_projection, VP, sizeof(VP));
    const float cot_y = 1.0f / std::tan(fov_y * 0.5f);
    out->fov_cotangent[0] = cot_y / aspect;
    out->fov_cotangent[1] = cot_y;
    out->depth_scale_bias[0] = -2.0f / (z_far - z_near);
    out->depth_scale_bias[1] = -(z_far + z_near) / (z_far - z_near);
    out->position[0] = eye[0];
    out->position[1] = eye[1];
    out->position[2] = eye[2];
    out->aspect = aspect;
    return GR_OK;
}

}  // extern "C"
