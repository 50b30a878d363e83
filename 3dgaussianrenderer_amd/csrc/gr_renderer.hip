// gr_renderer.hip -- object-level renderer behind the C-ABI: the device-side
// state of gauss_render::Renderer (core/renderer.h:23-49) and the frame of
// core/renderer.cpp:94-139 / app/main.cpp:326-433.
//
// Frame (one HIP stream, no host sync except the 4-byte K readback that sizes
// the tile list, as the reference's buildTileList does at
// core/gaussian_render.cu:581):
//   SH (degree > 0) -> project -> build count/scan, read K -> grow tile list
//   (grow-only, 1.25x headroom) -> emit -> radix sort -> ranges -> raster.
// The reference instead truncates a full tile list and doubles the capacity
// for the next frame (app/main.cpp:330-338,405); here a frame is never
// truncated.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gr_capi.h"
#include "gr_internal.hpp"
#include "gr_kernels.hpp"

using namespace gr;

namespace {

gr_status rfail(gr_status st, const std::string& msg);

#define R_HIP(call)                                                                           \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return rfail(e_ == hipErrorOutOfMemory ? GR_ERR_OUT_OF_MEMORY : GR_ERR_HIP,       \
                         std::string(#call) + " failed: " + hipGetErrorString(e_));           \
    } while (0)

#define R_CHECK(st)                  \
    do {                             \
        gr_status s_ = (st);         \
        if (s_ != GR_OK) return s_;  \
    } while (0)

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;
    gr_status ensure(size_t n, double headroom = 1.0) {
        if (n <= cap && p) return GR_OK;
        size_t want = std::max<size_t>((size_t)((double)n * headroom), std::max<size_t>(n, 1));
        if (p) R_HIP(hipFree(p));
        p = nullptr;
        cap = 0;
        R_HIP(hipMalloc((void**)&p, want * sizeof(T)));
        cap = want;
        return GR_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

}  // namespace

namespace {
gr_status rfail(gr_status st, const std::string& msg) {
    gr_detail::set_last_error(msg);
    return st;
}
}  // namespace

struct gr_renderer {
    int device = 0;
    gr_context* ctx = nullptr;
    hipEvent_t ev[7] = {};
    int32_t n = 0, degree = 0, sh_count = 3, precision = GR_PRECISION_FP32;
    DevBuf<float4> position, scale_rot, color, conic, ellipse;
    DevBuf<uint2> packed;        // precision form: scale + opacity, 16-bit
    DevBuf<uint16_t> sh16;       // precision form: aligned SH, 16-bit
    DevBuf<float2> clip_xy;
    DevBuf<float> clip_z, sh;
    DevBuf<int32_t> tile_range;
    DevBuf<uint64_t> keys[2];
    DevBuf<int32_t> values[2];
    DevBuf<uchar4> out;
    void* sort_temp = nullptr;
    size_t sort_temp_bytes = 0;
    int32_t sel = 0;
    int64_t last_k = 0;
    int32_t last_w = 0, last_h = 0;
};

extern "C" {

gr_status gr_renderer_create(gr_renderer** out, int32_t device, void* stream) {
    if (!out) return rfail(GR_ERR_INVALID_ARGUMENT, "gr_renderer_create: out is NULL");
    if (device >= 0) R_HIP(hipSetDevice(device));
    auto* r = new gr_renderer();
    R_HIP(hipGetDevice(&r->device));
    gr_status st = gr_context_create(&r->ctx, stream);
    if (st != GR_OK) {
        delete r;
        return st;
    }
    for (auto& e : r->ev) R_HIP(hipEventCreate(&e));
    *out = r;
    return GR_OK;
}

gr_status gr_renderer_destroy(gr_renderer* r) {
    if (!r) return GR_OK;
    (void)hipSetDevice(r->device);
    if (r->ctx) (void)gr_context_synchronize(r->ctx);
    r->position.release(); r->scale_rot.release(); r->color.release(); r->conic.release();
    r->ellipse.release(); r->clip_xy.release(); r->clip_z.release(); r->sh.release();
    r->packed.release(); r->sh16.release();
    r->tile_range.release(); r->keys[0].release(); r->keys[1].release();
    r->values[0].release(); r->values[1].release(); r->out.release();
    if (r->sort_temp) (void)hipFree(r->sort_temp);
    for (auto& e : r->ev) if (e) (void)hipEventDestroy(e);
    gr_context_destroy(r->ctx);
    delete r;
    return GR_OK;
}

gr_status gr_renderer_upload_model(gr_renderer* r, const gr_model_host* m) {
    if (!r || !m) return rfail(GR_ERR_INVALID_ARGUMENT, "gr_renderer_upload_model: NULL argument");
    if (m->num_gaussians < 0 || m->sh_degree < 0 || m->sh_degree > 3)
        return rfail(GR_ERR_INVALID_ARGUMENT, "model has a bad count or SH degree");
    if (m->num_gaussians > 0 && (!m->position || !m->scale_and_rotation))
        return rfail(GR_ERR_INVALID_ARGUMENT, "model position/scale arrays are NULL");
    if (m->sh_degree == 0 && m->num_gaussians > 0 && !m->color)
        return rfail(GR_ERR_INVALID_ARGUMENT, "degree-0 model needs the color array");
    if (m->sh_degree > 0 && m->num_gaussians > 0 && !m->sh)
        return rfail(GR_ERR_INVALID_ARGUMENT, "model SH array is NULL");
    if (m->precision < GR_PRECISION_FP32 || m->precision > GR_PRECISION_BF16)
        return rfail(GR_ERR_INVALID_ARGUMENT, "unknown precision");
    R_HIP(hipSetDevice(r->device));
    const int n = m->num_gaussians;
    r->n = n;
    r->degree = m->sh_degree;
    r->sh_count = 3 * (m->sh_degree + 1) * (m->sh_degree + 1);
    r->precision = m->precision;
    const size_t nn = (size_t)std::max(n, 1);
    R_CHECK(r->position.ensure(nn));
    R_CHECK(r->scale_rot.ensure(nn));
    R_CHECK(r->color.ensure(nn));
    R_CHECK(r->conic.ensure(nn));
    R_CHECK(r->ellipse.ensure(nn));
    R_CHECK(r->clip_xy.ensure(nn));
    R_CHECK(r->clip_z.ensure(nn));
    if (m->precision != GR_PRECISION_FP32) {
        // precision form: 16-bit SH and scale/opacity, quaternion bits in position.w
        const size_t groups = ((size_t)n + kShGroup - 1) / kShGroup;
        std::vector<float> pos((size_t)4 * nn);
        std::vector<uint16_t> pk((size_t)4 * nn);
        std::vector<uint16_t> sh16(m->sh_degree > 0 ? groups * kShGroup * (size_t)r->sh_count : 1);
        gr_status st = gr_pack_precision_model(m, m->precision, pos.data(), pk.data(), sh16.data());
        if (st != GR_OK) return st;
        R_CHECK(r->packed.ensure(nn));
        R_CHECK(r->sh16.ensure(sh16.size()));
        if (n > 0) {
            R_HIP(hipMemcpy(r->position.p, pos.data(), (size_t)n * 16, hipMemcpyHostToDevice));
            R_HIP(hipMemcpy(r->packed.p, pk.data(), (size_t)n * 8, hipMemcpyHostToDevice));
            if (m->sh_degree > 0)
                R_HIP(hipMemcpy(r->sh16.p, sh16.data(), sh16.size() * 2, hipMemcpyHostToDevice));
            else
                R_HIP(hipMemcpy(r->color.p, m->color, (size_t)n * 16, hipMemcpyHostToDevice));
        }
        return GR_OK;
    }
    if (n > 0) {
        R_HIP(hipMemcpy(r->position.p, m->position, (size_t)n * 16, hipMemcpyHostToDevice));
        R_HIP(hipMemcpy(r->scale_rot.p, m->scale_and_rotation, (size_t)n * 16, hipMemcpyHostToDevice));
        if (m->sh_degree == 0)
            R_HIP(hipMemcpy(r->color.p, m->color, (size_t)n * 16, hipMemcpyHostToDevice));
    }
    if (m->sh_degree > 0 && n > 0) {
        const size_t groups = ((size_t)n + kShGroup - 1) / kShGroup;
        const size_t total = groups * kShGroup * (size_t)r->sh_count;
        std::vector<float> aligned(total);
        gr_status st = gr_realign_spherical_harmonics(m->sh, aligned.data(), kShGroup, r->sh_count, n);
        if (st != GR_OK) return st;
        R_CHECK(r->sh.ensure(total));
        R_HIP(hipMemcpy(r->sh.p, aligned.data(), total * sizeof(float), hipMemcpyHostToDevice));
    }
    return GR_OK;
}

gr_status gr_renderer_render(gr_renderer* r, const gr_camera_data* cam, int32_t width, int32_t height,
                             uint8_t* out_rgba8, int32_t out_is_device, gr_render_stats* stats) {
    if (!r || !cam || !out_rgba8) return rfail(GR_ERR_INVALID_ARGUMENT, "gr_renderer_render: NULL argument");
    if (width <= 0 || height <= 0) return rfail(GR_ERR_INVALID_ARGUMENT, "bad frame size");
    R_HIP(hipSetDevice(r->device));
    gr_context* ctx = r->ctx;
    hipStream_t stream = ctx->stream;
    const FrameGeom geom = make_geom(width, height);
    const int tiles = geom.tiles_x * geom.tiles_y;
    R_CHECK(r->tile_range.ensure((size_t)2 * tiles));
    uchar4* outp = reinterpret_cast<uchar4*>(out_rgba8);
    if (!out_is_device) {
        R_CHECK(r->out.ensure((size_t)width * height));
        outp = r->out.p;
    }
    gr_global_args ga{};
    ga.camera_data = *cam;
    ga.position = reinterpret_cast<const float*>(r->position.p);
    ga.scale_and_rotation = reinterpret_cast<const float*>(r->scale_rot.p);
    ga.color = reinterpret_cast<float*>(r->color.p);
    ga.spherical_harmonics_degree = r->degree;
    ga.spherical_harmonics_count = r->sh_count;
    ga.spherical_harmonics = r->precision != GR_PRECISION_FP32 ? (const void*)r->sh16.p : (const void*)r->sh.p;
    ga.packed_scale_opacity = r->precision != GR_PRECISION_FP32 ? (const void*)r->packed.p : nullptr;
    ga.position_clip_xy = reinterpret_cast<float*>(r->clip_xy.p);
    ga.position_clip_z = r->clip_z.p;
    ga.conic = reinterpret_cast<float*>(r->conic.p);
    ga.screen_ellipse = reinterpret_cast<float*>(r->ellipse.p);
    ga.splat_count = r->n;
    ga.back_buffer = reinterpret_cast<uint8_t*>(outp);
    ga.tile_range = r->tile_range.p;
    ga.width = width;
    ga.height = height;
    ga.sh_precision = r->precision;
    R_CHECK(gr_set_global_args(ctx, &ga));

    gr_tile_list_args tl{};
    tl.keys = r->keys[0].p;
    tl.values = r->values[0].p;
    tl.capacity = (int32_t)std::min<size_t>(r->keys[0].cap, (size_t)0x7fffffff);
    R_CHECK(gr_set_tile_list_args(ctx, &tl));

    R_HIP(hipEventRecord(r->ev[0], stream));
    if (r->degree > 0) R_CHECK(gr_evaluate_spherical_harmonics(ctx, nullptr, r->n));
    R_HIP(hipEventRecord(r->ev[1], stream));
    R_CHECK(gr_evaluate_splat_clip_data(ctx, nullptr, r->n));
    R_HIP(hipEventRecord(r->ev[2], stream));

    int32_t size = 0;
    int64_t need = 0;
    R_CHECK(gr_build_tile_list(ctx, nullptr, 0, tl.capacity, &size, &need));
    if (need > tl.capacity) {
        if (need >= (int64_t)(1 << 30))
            return rfail(GR_ERR_CAPACITY, "tile list of " + std::to_string(need) + " entries exceeds 2^30");
        const size_t want = (size_t)need;
        R_CHECK(r->keys[0].ensure(want, 1.25));
        R_CHECK(r->keys[1].ensure(want, 1.25));
        R_CHECK(r->values[0].ensure(want, 1.25));
        R_CHECK(r->values[1].ensure(want, 1.25));
        tl.keys = r->keys[0].p;
        tl.values = r->values[0].p;
        tl.capacity = (int32_t)std::min<size_t>(r->keys[0].cap, (size_t)0x7fffffff);
        R_CHECK(gr_set_tile_list_args(ctx, &tl));
        R_CHECK(gr_build_tile_list(ctx, nullptr, 0, tl.capacity, &size, &need));
    }
    R_HIP(hipEventRecord(r->ev[3], stream));

    gr_double_buffer_u64 kb{{r->keys[0].p, r->keys[1].p}, 0};
    gr_double_buffer_i32 vb{{r->values[0].p, r->values[1].p}, 0};
    R_CHECK(gr_sort_tile_list(ctx, nullptr, size, &r->sort_temp, &r->sort_temp_bytes, &kb, &vb));
    R_HIP(hipEventRecord(r->ev[4], stream));
    r->sel = kb.selector;
    tl.keys = kb.buffers[kb.selector];
    tl.values = vb.buffers[vb.selector];
    tl.size = size;
    R_CHECK(gr_set_tile_list_args(ctx, &tl));
    R_CHECK(gr_evaluate_tile_range(ctx, nullptr, size));
    R_CHECK(gr_rasterize_tile(ctx, nullptr));
    R_HIP(hipEventRecord(r->ev[5], stream));
    if (!out_is_device)
        R_HIP(hipMemcpyAsync(out_rgba8, outp, (size_t)width * height * 4, hipMemcpyDeviceToHost, stream));
    R_HIP(hipEventRecord(r->ev[6], stream));
    R_HIP(hipEventSynchronize(r->ev[6]));
    r->last_k = need;
    r->last_w = width;
    r->last_h = height;
    if (stats) {
        float t[6];
        R_HIP(hipEventElapsedTime(&t[0], r->ev[0], r->ev[1]));
        R_HIP(hipEventElapsedTime(&t[1], r->ev[1], r->ev[2]));
        R_HIP(hipEventElapsedTime(&t[2], r->ev[2], r->ev[3]));
        R_HIP(hipEventElapsedTime(&t[3], r->ev[3], r->ev[4]));
        R_HIP(hipEventElapsedTime(&t[4], r->ev[4], r->ev[5]));
        R_HIP(hipEventElapsedTime(&t[5], r->ev[0], r->ev[5]));
        stats->sh_eval_ms = t[0];
        stats->project_ms = t[1];
        stats->build_tile_ms = t[2];
        stats->sort_ms = t[3];
        stats->raster_ms = t[4];
        stats->total_ms = t[5];
        stats->tile_list_size = need;
    }
    return GR_OK;
}

gr_status gr_renderer_frame_view(gr_renderer* r, gr_frame_view* v) {
    if (!r || !v) return rfail(GR_ERR_INVALID_ARGUMENT, "gr_renderer_frame_view: NULL argument");
    const FrameGeom g = make_geom(std::max(r->last_w, 1), std::max(r->last_h, 1));
    v->color = reinterpret_cast<const float*>(r->color.p);
    v->position_clip_xy = reinterpret_cast<const float*>(r->clip_xy.p);
    v->position_clip_z = r->clip_z.p;
    v->screen_ellipse = reinterpret_cast<const float*>(r->ellipse.p);
    v->conic = reinterpret_cast<const float*>(r->conic.p);
    v->sorted_keys = r->keys[r->sel].p;
    v->sorted_values = r->values[r->sel].p;
    v->tile_range = r->tile_range.p;
    v->tile_list_size = r->last_k;
    v->tiles_x = g.tiles_x;
    v->tiles_y = g.tiles_y;
    v->splat_count = r->n;
    return GR_OK;
}

}  // extern "C"
