// gr_renderer.hip -- object-level renderer behind the C-ABI: the device-side
// state of gauss_render::Renderer (core/renderer.h:23-49) and the frame of
// core/renderer.cpp:94-139 / app/main.cpp:326-433.
//
// Frame (one HIP stream; the only host sync is at the end, after the image):
//   front end, one pass per 256 Gaussians: SH colour (degree > 0), projection,
//      visibility (non-empty tile rect) and compaction of the block's visible
//      Gaussians into build records in the block's own staging slice
//   -> scan of the block counts (the record numbering) -> staging slot of each
//      record -> build count: per record its emit record (mask of passing
//      candidates, rect) and the depth sort's pair (depth word, record index
//      with the entry count packed above it)
//   -> 32-bit onesweep sort of the pairs -> the counts, now in depth order,
//      are scanned (K) into contiguous output offsets, with the window table
//      (the record holding every 512th entry)
//   -> emit (depth order, one wave per 512-entry window): the entries, as
//      8-byte (Gaussian index << 32 | tile id) words, in order, + tile digit
//      histograms
//   -> keys-only onesweep sort of the tile bits (bits [0, tile bits)), whose
//      last pass also writes the tile ranges
//   -> tile order (heaviest first) -> raster -> image and K readback.
// Entries sit in (depth word, Gaussian index) order before the last sort, so
// the stable tile-bit sort yields exactly the order of the reference's
// full-key sort of (tile << 32 | depth word) keys emitted in index order,
// with 2 passes over the K entries instead of 6 (the 4 depth passes run over
// the visible Gaussians only).
// The reference truncates a full tile list and doubles the capacity for the
// next frame (app/main.cpp:330-338,405); here a frame is never truncated:
// when K exceeds the list, the emit writes only what fits, the sort, ranges
// and raster see K > capacity and do nothing, and the frame is rendered
// again with a list grown to 1.25 K.
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

constexpr int kMaxMarks = 24;

struct gr_renderer {
    int device = 0;
    gr_context* ctx = nullptr;
    hipEvent_t ev[9] = {};          // per-stage events (only when stats are requested) + frame end
    hipEvent_t pass_ev[2] = {};     // around the tile-sort passes (GR_TIMING_SORT_PASSES / KERNELS, no stats)
    hipEvent_t front_ev[2] = {};    // around the front-end kernel (GR_TIMING_FRONT)
    bool front_timed = false;       // the last frame recorded front_ev
    hipEvent_t mark_ev[kMaxMarks] = {};   // per-kernel marks (timing level 2)
    int32_t timing = GR_TIMING_NONE;
    // per-kernel marks of the last frame: event handle and the name of the work before it
    int32_t marks = 0;
    hipEvent_t mark_at[kMaxMarks] = {};
    const char* mark_name[kMaxMarks] = {};
    hipEvent_t pass_begin = nullptr, pass_end = nullptr;   // events that bracketed the last passes
    int64_t last_visible = -1;
    int32_t last_passes = 0;
    int32_t n = 0, degree = 0, sh_count = 3, precision = GR_PRECISION_FP32;
    DevBuf<float4> position, scale_rot, color, conic;
    DevBuf<uint2> packed;        // precision form: scale + opacity, 16-bit
    DevBuf<uint16_t> sh16;       // precision form: aligned SH, 16-bit
    DevBuf<float2> clip_xy;
    DevBuf<float> clip_z, sh;
    DevBuf<int32_t> tile_range;
    DevBuf<uint64_t> entries[2];    // tile list: index << 32 | tile id (double buffer of the tile sort)
    DevBuf<uint64_t> full_keys;     // the list in the reference's form, made on inspection
    DevBuf<int32_t> full_values;
    DevBuf<uchar4> out;
    DevBuf<uint32_t> batch_counts;  // total K, scan scratch
    DevBuf<uint4> emit_recs;        // emit record per build record
    DevBuf<uint32_t> rec_prefix;    // output offset per depth-order position of the records
    DevBuf<uint32_t> dkeys[2];      // depth words of the visible Gaussians
    DevBuf<int32_t> dorder[2];      // their packed record words (record index | entry count)
    DevBuf<float4> records;         // staged build records of the visible Gaussians (2 x float4 each)
    DevBuf<uint32_t> blocks;        // front-end block counts (then offsets), the record total, scan scratch
    DevBuf<uint32_t> slots;         // staging slot of each record
    void* sort_temp = nullptr;      // the tile-bit sort's scratch
    size_t sort_temp_bytes = 0;
    void* dsort_temp = nullptr;     // the depth sort's scratch (both are zeroed by the front end)
    size_t dsort_temp_bytes = 0;
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
    // Timing events inside the frame order nothing for the host, so they skip the system-scope
    // cache fence (each default event costs about 6 us of idle GPU between the stages it
    // separates); the last one, after the image and K copies, keeps it: the host reads those.
    for (int k = 0; k < 8; ++k) R_HIP(hipEventCreateWithFlags(&r->ev[k], hipEventDisableSystemFence));
    R_HIP(hipEventCreate(&r->ev[8]));
    for (auto& e : r->pass_ev) R_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    for (auto& e : r->front_ev) R_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    for (auto& e : r->mark_ev) R_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    *out = r;
    return GR_OK;
}

gr_status gr_renderer_destroy(gr_renderer* r) {
    if (!r) return GR_OK;
    (void)hipSetDevice(r->device);
    if (r->ctx) (void)gr_context_synchronize(r->ctx);
    r->position.release(); r->scale_rot.release(); r->color.release(); r->conic.release();
    r->clip_xy.release(); r->clip_z.release(); r->sh.release();
    r->packed.release(); r->sh16.release();
    r->tile_range.release(); r->entries[0].release(); r->entries[1].release(); r->full_keys.release();
    r->full_values.release(); r->out.release();
    r->batch_counts.release();
    r->dkeys[0].release(); r->dkeys[1].release(); r->dorder[0].release(); r->dorder[1].release();
    r->rec_prefix.release(); r->emit_recs.release();
    r->records.release(); r->blocks.release(); r->slots.release();
    if (r->sort_temp) (void)hipFree(r->sort_temp);
    if (r->dsort_temp) (void)hipFree(r->dsort_temp);
    for (auto& e : r->ev) if (e) (void)hipEventDestroy(e);
    for (auto& e : r->pass_ev) if (e) (void)hipEventDestroy(e);
    for (auto& e : r->front_ev) if (e) (void)hipEventDestroy(e);
    for (auto& e : r->mark_ev) if (e) (void)hipEventDestroy(e);
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

static gr_status ensure_temp(void*& p, size_t& have, size_t need) {
    if (p && have >= need) return GR_OK;
    if (p) R_HIP(hipFree(p));
    p = nullptr;
    have = 0;
    const size_t want = need + need / 4;
    R_HIP(hipMalloc(&p, want));
    have = want;
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
    if (tiles >= (1 << 20)) return rfail(GR_ERR_INVALID_ARGUMENT, "frame has too many tiles");
    R_CHECK(r->tile_range.ensure((size_t)3 * tiles));   // ranges, then the raster's tile order
    uchar4* outp = reinterpret_cast<uchar4*>(out_rgba8);
    if (!out_is_device) {
        R_CHECK(r->out.ensure((size_t)width * height));
        outp = r->out.p;
    }
    const int n = r->n;
    const int batches = (n + kBuildBatch - 1) / kBuildBatch;
    const size_t nn = (size_t)std::max(n, 1);
    R_CHECK(r->batch_counts.ensure(2 + scan_scratch_words((int32_t)nn)));
    R_CHECK(r->emit_recs.ensure(nn));
    R_CHECK(r->rec_prefix.ensure(nn));
    for (int k = 0; k < 2; ++k) {
        R_CHECK(r->entries[k].ensure(1));
        R_CHECK(r->dkeys[k].ensure(nn));
        R_CHECK(r->dorder[k].ensure(nn));
    }
    const int nb = front_blocks(n);   // 0 for an empty scene
    R_CHECK(r->records.ensure((size_t)2 * kFrontThreads * std::max(nb, 1)));
    R_CHECK(r->blocks.ensure((size_t)nb + 1 + scan_scratch_words(nb)));
    R_CHECK(r->slots.ensure(nn));
    // one scratch serves the depth sort and the tile-bit sort; sized before anything is enqueued
    const int32_t capacity = (int32_t)std::min<size_t>(std::min(r->entries[0].cap, r->entries[1].cap), (1u << 30) - 1);
    R_CHECK(ensure_temp(r->sort_temp, r->sort_temp_bytes, radix_sort_temp_bytes(capacity, 0, 32)));
    R_CHECK(ensure_temp(r->dsort_temp, r->dsort_temp_bytes, radix_sort_temp_bytes((int32_t)nn, 0, 32)));
    // what the frame needs set before it runs -- the tile ranges to -1 (tiles without entries) and
    // both sorts' histograms, tickets and first look-back regions to 0 -- the front end sets in
    // passing (no fill launches)
    FillSpan dzero[2], tzero[2];
    uint32_t* dhist = nullptr;
    uint32_t* thist = nullptr;
    R_HIP(radix_sort_zero_spans((int32_t)nn, r->dsort_temp, r->dsort_temp_bytes, dzero, &dhist));
    R_HIP(radix_sort_zero_spans(capacity, r->sort_temp, r->sort_temp_bytes, tzero, &thist));

    ViewDesc vd{};
    vd.cam = camera_args(*cam);
    vd.geom = geom;
    vd.tile_base = 0;
    vd.tiles = tiles;
    vd.out = outp;
    uint32_t* k_total = r->batch_counts.p;

    // Timing.  Per-stage events run only when stats are requested; per-kernel marks only at
    // GR_TIMING_KERNELS; the tile-sort passes are bracketed at GR_TIMING_SORT_PASSES and up (and
    // by the stage events, which share them).  Each event leaves the GPU idle for a few
    // microseconds, so a frame asked for none of them records none.
    const bool stages = stats != nullptr;
    const bool kmarks = r->timing == GR_TIMING_KERNELS;
    const bool front_ev = r->timing == GR_TIMING_FRONT;
    r->front_timed = front_ev;
    r->marks = 0;
    auto stage = [&](int k) -> hipError_t { return stages ? hipEventRecord(r->ev[k], stream) : hipSuccess; };
    auto mark = [&](const char* name) -> hipError_t {
        if (!kmarks || r->marks >= kMaxMarks) return hipSuccess;
        r->mark_at[r->marks] = r->mark_ev[r->marks];
        r->mark_name[r->marks] = name;
        return hipEventRecord(r->mark_ev[r->marks++], stream);
    };
    uint32_t* vis_total = r->blocks.p + nb;   // visible Gaussians (records)
    R_HIP(stage(0));
    // the second list buffer (capacity * 2 words) holds, until the tile sort's first pass, the
    // emit's window table and after it the count pass's listed tiles of footprints over 64
    // candidates
    uint32_t* const list1 = reinterpret_cast<uint32_t*>(r->entries[1].p);
    const int64_t side_at = (emit_window_words(capacity) + 63) / 64 * 64;
    const int64_t side_words = std::min<int64_t>(2 * (int64_t)capacity - side_at, UINT32_MAX);
    R_HIP(mark("begin"));
    if (front_ev) R_HIP(hipEventRecord(r->front_ev[0], stream));
    {
        // SH colour + projection + compaction of the visible Gaussians into build records
        FrontArgs a{};
        a.fills[0] = {reinterpret_cast<uint32_t*>(r->tile_range.p), 2 * (int64_t)tiles, 0xffffffffu};
        a.fills[1] = dzero[0];
        a.fills[2] = dzero[1];
        a.fills[3] = tzero[0];
        a.fills[4] = tzero[1];
        a.nfills = 5;
        a.position = r->position.p;
        a.scale_rotation = r->scale_rot.p;
        a.packed = r->packed.p;
        a.sh = r->precision != GR_PRECISION_FP32 ? (const void*)r->sh16.p : (const void*)r->sh.p;
        a.degree = n > 0 ? r->degree : 0;
        a.sh_count = r->sh_count;
        a.precision = r->precision;
        a.color = r->color.p;
        a.clip_xy = r->clip_xy.p;
        a.clip_z = r->clip_z.p;
        a.ellipse = nullptr;   // one view: the ellipses live in the records
        a.conic = r->conic.p;
        a.records = r->records.p;
        a.block_counts = r->blocks.p;
        a.count = n;
        a.nviews = 1;
        a.v0 = vd;
        R_HIP(launch_front(a, stream));
    }
    if (front_ev) R_HIP(hipEventRecord(r->front_ev[1], stream));
    R_HIP(stage(1));
    R_HIP(mark("front"));
    // record numbering: block offsets + the total, then each record's staging slot
    R_HIP(launch_build_scan(r->blocks.p, nb, vis_total, r->blocks.p + nb + 1, stream));
    R_HIP(launch_record_slots(r->blocks.p, vis_total, nb, kFrontThreads, r->slots.p, stream));
    R_HIP(mark("numbering"));
    // emit record and depth pair of every record
    BuildArgs b{};
    b.records = r->records.p;
    b.record_slots = r->slots.p;
    b.record_total = vis_total;
    b.depth_keys = r->dkeys[0].p;
    b.depth_words = reinterpret_cast<uint32_t*>(r->dorder[0].p);
    b.emit_records = r->emit_recs.p;
    b.index_bits = packed_index_bits(std::max(n, 1));
    b.entries = r->entries[0].p;
    if (side_words > 0) {
        b.side = list1 + side_at;
        b.side_words = side_words;
    }
    b.count = n;
    b.capacity = capacity;
    b.batches = batches;
    b.nviews = 1;
    b.v0 = vd;
    R_HIP(launch_build_count(b, stream));
    R_HIP(stage(2));
    R_HIP(stage(3));
    R_HIP(mark("record_count"));
    // depth order of the records
    int32_t dsel = 0;
    {
        uint32_t* dk[2] = {r->dkeys[0].p, r->dkeys[1].p};
        int32_t* dv[2] = {r->dorder[0].p, r->dorder[1].p};
        R_HIP(radix_sort_pairs_u32(dk, dv, &dsel, n, vis_total, 0, 32, r->dsort_temp, r->dsort_temp_bytes, stream,
                                   /*zeroed=*/true));
    }
    R_HIP(stage(4));
    R_HIP(mark("depth_sort"));
    // output offsets: the counts carried by the sorted words, scanned in depth order (total K),
    // and the emit windows' first records.  The window table lives in the second list buffer,
    // which the tile sort's first pass overwrites only after the emit (capacity / kEmitWindow + 1 words
    // in capacity * 2 words).
    const uint32_t* sorted_words = reinterpret_cast<const uint32_t*>(r->dorder[dsel].p);
    uint32_t* windows = list1;
    R_HIP(launch_build_scan_packed(sorted_words, n, vis_total, r->emit_recs.p, b.index_bits, k_total, k_total + 1,
                                   r->rec_prefix.p, windows, emit_window_words(capacity), stream));
    R_HIP(mark("offsets"));
    b.sorted_words = sorted_words;
    b.depth_prefix = r->rec_prefix.p;
    b.windows = windows;
    b.entry_total = k_total;
    const int tile_bits = tile_bits_for(tiles);
    b.tile_passes = (tile_bits + 7) / 8;
    b.tile_hist = b.tile_passes ? thist : nullptr;   // zeroed by the front end; one tile: no pass
    b.tile_digit_bits = sort_digit_bits(0, tile_bits);
    R_HIP(launch_build_emit(b, stream));
    R_HIP(mark("emit"));
    // tile bits: the sort records its begin event after its histogram scan and its end event
    // after its passes, so the interval is the passes alone (the stage split shares the two)
    int32_t sel = 0;
    {
        hipEvent_t pb = nullptr, pe = nullptr;
        if (stages) {
            pb = r->ev[5];
            pe = r->ev[6];
        } else if (r->timing == GR_TIMING_SORT_PASSES || r->timing == GR_TIMING_KERNELS) {
            pb = r->pass_ev[0];
            pe = r->pass_ev[1];
        }
        uint64_t* kb[2] = {r->entries[0].p, r->entries[1].p};
        // the last pass writes the tile ranges; tiles without entries keep the -1 the front end set
        RangeArgs ra{};
        ra.tile_shift = 0;
        ra.tile_range = r->tile_range.p;
        ra.tiles = tiles;
        R_HIP(radix_sort_keys_prepared(kb, &sel, capacity, k_total, 0, tile_bits, r->sort_temp, r->sort_temp_bytes,
                                       stream, pb, pe, &ra));
        r->last_passes = capacity > 1 ? (tile_bits + 7) / 8 : 0;
        if (r->last_passes == 0 && pb) {   // no passes ran: the events still bracket the (empty) sort
            R_HIP(hipEventRecord(pb, stream));
            R_HIP(hipEventRecord(pe, stream));
        }
        r->pass_begin = pb;
        r->pass_end = pe;
        if (kmarks && pb && r->marks + 2 <= kMaxMarks) {
            // the pass events double as marks: histogram scan, then the passes
            r->mark_at[r->marks] = pb;
            r->mark_name[r->marks++] = "tile_sort_scan";
            r->mark_at[r->marks] = pe;
            r->mark_name[r->marks++] = "tile_sort_passes";
        }
    }
    r->sel = sel;
    {
        RasterArgs a{};
        a.tile_range = r->tile_range.p;
        // heaviest tiles first (at 6M/1080p, raster 218 -> 210 us with the order kernel)
        uint32_t* order = reinterpret_cast<uint32_t*>(r->tile_range.p + 2 * (size_t)tiles);
        R_HIP(launch_tile_order(r->tile_range.p, tiles, order, stream));
        a.order = order;
        a.entries = r->entries[sel].p;
        a.clip_xy = r->clip_xy.p;
        a.conic = r->conic.p;
        a.color = r->color.p;
        a.count = n;
        a.tiles = tiles;
        a.nviews = 1;
        a.v0 = vd;
        R_HIP(launch_raster(a, stream));
    }
    R_HIP(stage(7));
    R_HIP(mark("raster"));
    if (!out_is_device)
        R_HIP(hipMemcpyAsync(out_rgba8, outp, (size_t)width * height * 4, hipMemcpyDeviceToHost, stream));
    R_HIP(hipMemcpyAsync(ctx->pinned, k_total, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    if (kmarks) R_HIP(hipMemcpyAsync(ctx->pinned + 1, vis_total, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    R_HIP(hipEventRecord(r->ev[8], stream));
    R_HIP(hipEventSynchronize(r->ev[8]));
    const int64_t need = (int64_t)ctx->pinned[0];
    if (need > capacity) {
        // the list was too small: nothing was sorted or drawn; grow it and render the frame again
        if (need >= (int64_t)(1 << 30))
            return rfail(GR_ERR_CAPACITY, "tile list of " + std::to_string(need) + " entries exceeds 2^30");
        for (int k = 0; k < 2; ++k) R_CHECK(r->entries[k].ensure((size_t)need, 1.25));
        return gr_renderer_render(r, cam, width, height, out_rgba8, out_is_device, stats);
    }
    r->last_k = need;
    r->last_visible = kmarks ? (int64_t)ctx->pinned[1] : -1;
    r->last_w = width;
    r->last_h = height;
    if (stats) {
        // events: 0 front end (SH + projection + compaction) 1 count 2,3 depth sort 4 offsets +
        // emit + the tile sort's histogram scan 5 tile-bit sort passes (the last one writes the
        // tile ranges) 6 raster 7
        float t[8];
        for (int k = 0; k < 7; ++k) R_HIP(hipEventElapsedTime(&t[k], r->ev[k], r->ev[k + 1]));
        R_HIP(hipEventElapsedTime(&t[7], r->ev[0], r->ev[7]));
        stats->sh_eval_ms = t[0];   // the fused front end: SH, projection and compaction
        stats->project_ms = 0.0;    // folded into the front end
        stats->build_tile_ms = t[1] + t[4];   // numbering, count, offsets, emit (+ histogram scan)
        stats->sort_ms = t[3] + t[5];         // depth sort + tile-bit sort (with the tile ranges)
        stats->raster_ms = t[6];              // raster
        stats->total_ms = t[7];
        stats->tile_list_size = need;
    }
    return GR_OK;
}

gr_status gr_renderer_set_timing(gr_renderer* r, int32_t level) {
    if (!r || level < GR_TIMING_NONE || level > GR_TIMING_FRONT)
        return rfail(GR_ERR_INVALID_ARGUMENT, "gr_renderer_set_timing: bad argument");
    r->timing = level;
    return GR_OK;
}

gr_status gr_renderer_kernel_times(gr_renderer* r, gr_kernel_times* out) {
    if (!r || !out) return rfail(GR_ERR_INVALID_ARGUMENT, "gr_renderer_kernel_times: NULL argument");
    *out = gr_kernel_times{};
    out->tile_sort_passes = r->last_passes;
    out->tile_list_size = r->last_k;
    out->visible = r->last_visible;
    if (r->last_passes > 0 && r->pass_begin && r->pass_end) {
        float ms = 0.0f;
        R_HIP(hipEventElapsedTime(&ms, r->pass_begin, r->pass_end));
        out->tile_sort_pass_ms = ms;
    } else {
        out->tile_sort_pass_ms = -1.0;
    }
    out->front_ms = -1.0;
    if (r->front_timed) {
        float ms = 0.0f;
        R_HIP(hipEventElapsedTime(&ms, r->front_ev[0], r->front_ev[1]));
        out->front_ms = ms;
    }
    // mark k closes the work named mark_name[k] that began at mark k - 1
    out->kernels = 0;
    for (int k = 1; k < r->marks && out->kernels < GR_KERNEL_TIMES_MAX; ++k) {
        float ms = 0.0f;
        R_HIP(hipEventElapsedTime(&ms, r->mark_at[k - 1], r->mark_at[k]));
        const int i = out->kernels++;
        out->kernel_ms[i] = ms;
        std::snprintf(out->kernel_name[i], sizeof(out->kernel_name[i]), "%s", r->mark_name[k]);
    }
    return GR_OK;
}

gr_status gr_renderer_frame_view(gr_renderer* r, gr_frame_view* v) {
    if (!r || !v) return rfail(GR_ERR_INVALID_ARGUMENT, "gr_renderer_frame_view: NULL argument");
    const FrameGeom g = make_geom(std::max(r->last_w, 1), std::max(r->last_h, 1));
    v->color = reinterpret_cast<const float*>(r->color.p);
    v->position_clip_xy = reinterpret_cast<const float*>(r->clip_xy.p);
    v->position_clip_z = r->clip_z.p;
    v->screen_ellipse = nullptr;   // the frame keeps the ellipses in its build records only
    v->conic = reinterpret_cast<const float*>(r->conic.p);
    // the list is held as (index << 32 | tile) entries in (tile, depth word, index) order; its
    // reference-form keys and values are made here, once per frame inspected
    const int64_t k = r->last_k;
    if (k > 0) {
        R_CHECK(r->full_keys.ensure((size_t)k));
        R_CHECK(r->full_values.ensure((size_t)k));
        R_HIP(launch_expand_keys(r->entries[r->sel].p, r->clip_z.p, r->full_keys.p, r->full_values.p, k,
                                 r->ctx->stream));
        R_HIP(hipStreamSynchronize(r->ctx->stream));
    }
    v->sorted_keys = k > 0 ? r->full_keys.p : nullptr;
    v->sorted_values = k > 0 ? r->full_values.p : nullptr;
    v->tile_range = r->tile_range.p;
    v->tile_list_size = r->last_k;
    v->tiles_x = g.tiles_x;
    v->tiles_y = g.tiles_y;
    v->splat_count = r->n;
    return GR_OK;
}

}  // extern "C"
