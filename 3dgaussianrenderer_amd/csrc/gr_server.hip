// gr_server.hip -- continuous-batching render server over resident scenes
// (include/gr_capi.h, "serving").  No counterpart in the reference, which
// renders one scene one frame at a time (app/main.cpp:326-472).
//
// Threads: callers submit from any thread into a locked queue; one scheduler
// thread owns every HIP call of the render path.  Each slot is a stream plus
// a grow-only workspace; a batch (up to max_views requests of one scene, any
// mix of resolutions) runs on one slot in two phases:
//   A  camera block H2D -> front end, one pass over the scene for all views:
//      SH colour, projection, visibility and compaction of every view's
//      visible Gaussians into build records (each block of the scene into its own staging
//      slice) -> record numbering (scan of the block counts, staging slot per record) ->
//      build count per record (entries, mask of passing candidates, depth word) -> one
//      32-bit sort of all views' records by depth
//      word -> output offsets from the counts in depth order -> K readback
//   B  emit (entries at their depth-order offsets, tile-digit histograms)
//      -> onesweep sort of the tile bits (device-side K) -> ranges -> raster
//      (all views) -> images D2H into pinned staging
// Tile ids are global over the batch's views, so one sort serves all views,
// and within a tile entries keep (depth word, index) order.
// Between A and B the host reads K and grows the slot's tile-list buffers
// within the HBM budget (exact admission).  A phase whose (scene, batch
// shape, workspace) repeats is captured as a hipGraph and replayed from then
// on (a shape seen once runs eagerly: capture and instantiation cost more
// host time than the launches they save); the camera block is the only
// per-batch input and travels through a memcpy node from pinned memory.
// Several slots run concurrently, so batches of different scenes (or of one
// scene) overlap on the device.
// Copy-out threads move finished images from a slot's pinned staging into
// the callers' buffers and post the completions, in chunks spread over the
// threads; each slot has two staging buffers, so its next batch runs while
// the last one is copied out.  The scheduler thread never touches image
// bytes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/gr_capi.h"
#include "gr_internal.hpp"
#include "gr_kernels.hpp"

using namespace gr;

namespace {

gr_status sfail(gr_status st, const std::string& msg) {
    gr_detail::set_last_error(msg);
    return st;
}

#define S_HIP(call)                                                                          \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return sfail(e_ == hipErrorOutOfMemory ? GR_ERR_OUT_OF_MEMORY : GR_ERR_HIP,      \
                         std::string(#call) + " failed: " + hipGetErrorString(e_));          \
    } while (0)

#define S_CHECK(st)                  \
    do {                             \
        gr_status s_ = (st);         \
        if (s_ != GR_OK) return s_;  \
    } while (0)

// device allocation with byte accounting; workspaces grow with 25 % headroom, resident
// scene buffers (exact = true) take exactly the bytes admitted for them
struct Dev {
    void* p = nullptr;
    size_t bytes = 0;
    gr_status ensure(size_t need, std::atomic<int64_t>& acct, int* grows = nullptr, bool exact = false) {
        if (need <= bytes && p) return GR_OK;
        const size_t want = exact ? std::max<size_t>(need, 1) : std::max<size_t>(need + need / 4, 256);
        if (p) {
            S_HIP(hipFree(p));
            acct -= (int64_t)bytes;
        }
        p = nullptr;
        bytes = 0;
        S_HIP(hipMalloc(&p, want));
        bytes = want;
        acct += (int64_t)want;
        if (grows) ++*grows;
        return GR_OK;
    }
    // Net change of the accounted bytes if ensure(need) ran now (the old block is freed first).
    int64_t growth(size_t need) const {
        if (need <= bytes && p) return 0;
        return (int64_t)std::max<size_t>(need + need / 4, 256) - (int64_t)bytes;
    }
    void release(std::atomic<int64_t>& acct) {
        if (p) {
            (void)hipFree(p);
            acct -= (int64_t)bytes;
        }
        p = nullptr;
        bytes = 0;
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

struct Pinned {
    void* p = nullptr;
    size_t bytes = 0;
    gr_status ensure(size_t need, bool* grew = nullptr) {
        if (need <= bytes && p) return GR_OK;
        const size_t want = std::max<size_t>(need + need / 4, 256);
        if (p) S_HIP(hipHostFree(p));
        p = nullptr;
        bytes = 0;
        S_HIP(hipHostMalloc(&p, want, hipHostMallocDefault));
        bytes = want;
        if (grew) *grew = true;
        return GR_OK;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

struct Scene {
    int32_t id = -1;
    int32_t n = 0, degree = 0, sh_count = 3, precision = 0;
    Dev position, scale_rot, packed, sh, color0;
    int64_t bytes = 0;
    int inflight = 0;
    bool removing = false;
    void release(std::atomic<int64_t>& acct) {
        for (Dev* d : {&position, &scale_rot, &packed, &sh, &color0}) d->release(acct);
    }
};

struct Pending {
    uint64_t id;
    gr_request req;
    double submit_ms;
};

enum class SlotState { Idle, PhaseA, WaitMemory, PhaseB };

struct Workspace {
    Dev color, xy, con;              // [V][N]
    Dev blocks;                       // front-end block counts (then offsets), the record count, scan scratch
    Dev dkeys0, dkeys1, dorder0, dorder1;   // [<= V*N] depth words and packed record words of visible Gaussians
    Dev records;                      // [blocks][256 * V][2] float4 staged build records
    Dev slots;                        // [<= V*N] staging slot of each record
    Dev dsort_temp;                   // depth sort scratch
    Dev counts;                       // K (device) + scan scratch
    Dev emit_recs;                    // [V*N] emit record per record
    Dev rec_prefix;                   // [V*N] output offset per depth-order position of the records
    Dev list0, list1;                 // tile lists: Gaussian index << 32 | global tile id
    Dev sort_temp;
    Dev ranges;                       // 2 * tiles
    Dev images;                       // sum of W*H*4
    Dev views;                        // ViewDesc[V]
    Pinned views_host, k_host;
    Pinned img_host[2];               // image staging, alternating between batches
    uint64_t generation = 0;          // bumps when a pointer a captured graph holds changes
    std::vector<Dev*> devs() {
        return {&color, &xy, &con, &blocks, &dkeys0, &dkeys1, &dorder0, &dorder1, &records, &slots,
                &emit_recs, &rec_prefix,
                &dsort_temp, &counts, &list0, &list1, &sort_temp, &ranges, &images, &views};
    }
};

struct GraphKey {
    int32_t scene;
    uint64_t gen;
    std::vector<int32_t> shape;       // V, then (W, H) per view, then capacity for phase B
    bool operator<(const GraphKey& o) const {
        return std::tie(scene, gen, shape) < std::tie(o.scene, o.gen, o.shape);
    }
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    Workspace ws;
    SlotState state = SlotState::Idle;
    std::vector<Pending> batch;
    int32_t scene = -1;
    int32_t tiles_total = 0;
    int32_t tile_bits = 1;            // bits of the batch's global tile ids (the tile sort's key)
    int64_t img_bytes = 0;
    uint32_t k = 0;
    uint32_t capacity = 0;
    int32_t dsel = 0;                 // depth-sort buffer holding the sorted record words
    double start_ms = 0.0;
    int img_cur = 0;                  // staging buffer of the batch in flight
    std::map<GraphKey, hipGraphExec_t> graphs;
    std::map<GraphKey, int> seen;     // phase shapes run eagerly so far
};

// One view's copy-out: the chunks still to copy and the completion posted after the last.
struct ViewCopy {
    gr_completion c{};
    int32_t scene = -1;
    std::atomic<int> left{0};
    std::atomic<int>* staging_busy = nullptr;   // views of the staging buffer still copying
};

struct CopyChunk {
    const uint8_t* src;
    uint8_t* dst;
    size_t bytes;
    std::shared_ptr<ViewCopy> view;
};

#ifndef GR_COPY_THREADS
#define GR_COPY_THREADS 4   // build-time override for sweeps
#endif
constexpr int kCopyThreads = GR_COPY_THREADS;
constexpr size_t kCopyChunk = size_t(2) << 20;

}  // namespace

struct gr_server {
    gr_server_config cfg{};
    int device = 0;
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_submit, cv_done;
    std::deque<Pending> incoming;
    std::deque<gr_completion> done;
    std::map<int32_t, Scene*> scenes;
    std::atomic<int32_t> largest_n{0};               // most Gaussians of any scene added so far
    int32_t next_scene = 0;
    uint64_t next_request = 1;
    std::atomic<int64_t> resident{0}, workspace{0};
    std::atomic<int64_t> peak{0}, batches{0}, views{0}, graph_launches{0}, budget_waits{0};
    std::atomic<int64_t> graph_captures{0};
    std::atomic<int64_t> capture_us{0}, copy_us{0};
    std::atomic<int64_t> reallocs{0}, realloc_us{0};
    std::vector<Slot> slots;
    std::unique_ptr<std::atomic<int>[]> staging_busy;   // [slot][2]
    std::mutex copy_mu;
    std::condition_variable copy_cv;
    std::deque<CopyChunk> copy_q;
    std::vector<std::thread> copiers;
    bool copy_stop = false;                          // under copy_mu
    std::map<int32_t, std::deque<Pending>> queues;   // scheduler-owned
    std::map<int32_t, int> batch_limit;              // scheduler-owned: one-shot cap on a scene's next batch
    std::thread worker;
    std::atomic<bool> stop{false};
    std::string fatal;
    // Admission is check-then-allocate: the scheduler (workspaces) and add_scene callers
    // (resident scenes) each hold this lock from the budget check through the allocations it
    // admits, so concurrent admissions cannot both pass against the same headroom.
    std::mutex acct_mu;
    // message of the last request that completed with an error (the scheduler thread's
    // failures never reach the submitting thread's gr_last_error)
    std::mutex err_mu;
    std::string last_error;

    double now_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    int64_t used() const { return resident.load() + workspace.load(); }
    void note_peak() {
        const int64_t u = used();
        int64_t p = peak.load();
        while (u > p && !peak.compare_exchange_weak(p, u)) {
        }
    }
    // `n` buffers were (re)allocated since t0
    void note_realloc(int n, std::chrono::steady_clock::time_point t0) {
        if (n <= 0) return;
        reallocs += n;
        realloc_us += (int64_t)std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
    }
    bool over_budget(int64_t extra) const {
        return cfg.hbm_budget_bytes > 0 && used() + extra > cfg.hbm_budget_bytes;
    }
};

namespace {

constexpr int32_t kBatchTiles = 1 << 16;   // tiles per batch beyond which views go to the next batch
constexpr int64_t kMaxRecords = (1 << 30) - 1;   // views x Gaussians of a batch (depth sort's count)

// Views of an n-Gaussian scene one batch may hold: each view adds n depth-sort records.
// At most kFrontMaxViews views per batch (the front end's one pass over the scene).
int max_views_for(const gr_server* s, int32_t n) {
    const int64_t cap = std::min<int64_t>(std::min<int64_t>(s->cfg.max_views_per_batch, kFrontMaxViews),
                                          kMaxRecords / std::max(n, 1));
    return (int)std::max<int64_t>(1, cap);
}

Scene* scene_ptr(gr_server* s, int32_t id) {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->scenes.find(id);
    return it == s->scenes.end() ? nullptr : it->second;
}

void complete(gr_server* s, Slot& sl, int32_t status) {
    const double t = s->now_ms();
    if (status != GR_OK) {
        // the failing step left its message on this (the scheduler's) thread
        std::lock_guard<std::mutex> lk(s->err_mu);
        s->last_error = gr_detail::get_last_error();
    }
    std::lock_guard<std::mutex> lk(s->mu);
    for (const Pending& p : sl.batch) {
        gr_completion c{};
        c.request_id = p.id;
        c.user_tag = p.req.user_tag;
        c.status = status;
        c.batch_views = (int32_t)sl.batch.size();
        c.submit_ms = p.submit_ms;
        c.start_ms = sl.start_ms;
        c.finish_ms = t;
        c.tile_list_size = sl.k;
        s->done.push_back(c);
    }
    auto it = s->scenes.find(sl.scene);
    if (it != s->scenes.end()) it->second->inflight -= (int)sl.batch.size();
    s->cv_done.notify_all();
}

// Posts a view whose image has been copied out; its staging buffer may be reused afterwards.
void post_view(gr_server* s, ViewCopy& v) {
    v.c.finish_ms = s->now_ms();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->done.push_back(v.c);
        auto it = s->scenes.find(v.scene);
        if (it != s->scenes.end()) it->second->inflight -= 1;
    }
    v.staging_busy->fetch_sub(1);
    s->cv_done.notify_all();
}

// Copy-out thread: chunks in submission order until stopped with an empty queue.
void copier(gr_server* s) {
    for (;;) {
        CopyChunk ch;
        {
            std::unique_lock<std::mutex> lk(s->copy_mu);
            s->copy_cv.wait(lk, [&] { return s->copy_stop || !s->copy_q.empty(); });
            if (s->copy_q.empty()) return;
            ch = std::move(s->copy_q.front());
            s->copy_q.pop_front();
        }
        const auto t = std::chrono::steady_clock::now();
        std::memcpy(ch.dst, ch.src, ch.bytes);
        s->copy_us += (int64_t)std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t).count();
        if (ch.view->left.fetch_sub(1) == 1) post_view(s, *ch.view);
    }
}

// Phase-A buffers of a slot and the bytes a batch needs in each.
struct PhaseANeed {
    Dev* dev;
    int64_t bytes;
};

std::vector<PhaseANeed> phase_a_needs(Slot& sl, const Scene& sc, size_t nviews, int64_t img_bytes,
                                      int32_t tiles_total) {
    Workspace& w = sl.ws;
    const int64_t V = (int64_t)nviews, N = sc.n, nb = front_blocks(sc.n);
    return {{&w.color, V * N * 16},
            {&w.xy, V * N * 8},
            {&w.con, V * N * 16},
            {&w.blocks, 4 * (nb + 1 + scan_scratch_words((int32_t)nb))},
            {&w.dkeys0, V * N * 4},
            {&w.dkeys1, V * N * 4},
            {&w.dorder0, V * N * 4},
            {&w.dorder1, V * N * 4},
            {&w.records, V * std::max<int64_t>(nb, 1) * kFrontThreads * 32},
            {&w.slots, V * N * 4},
            {&w.dsort_temp, (int64_t)radix_sort_temp_bytes((int32_t)std::max<int64_t>(V * N, 1), 0, 32)},
            {&w.counts, 4 * (2 + scan_scratch_words((int32_t)(V * N)))},
            {&w.emit_recs, V * N * 16},
            {&w.rec_prefix, V * N * 4},
            {&w.images, img_bytes},
            {&w.ranges, 12 * (int64_t)tiles_total},   // ranges, then the raster's tile order
            {&w.views, V * (int64_t)sizeof(ViewDesc)}};
}

// Bytes phase A would add to the slot's workspace for this batch.
int64_t phase_a_growth(Slot& sl, const Scene& sc, size_t nviews, int64_t img_bytes, int32_t tiles_total) {
    int64_t g = 0;
    for (const PhaseANeed& n : phase_a_needs(sl, sc, nviews, img_bytes, tiles_total)) g += n.dev->growth((size_t)n.bytes);
    return g;
}

int64_t fresh_bytes(int64_t need) { return (int64_t)std::max<size_t>((size_t)need + (size_t)need / 4, 256); }

// Device bytes of a workspace allocated from nothing for the current batch with `entries` tile-list entries.
int64_t fresh_workspace_bytes(Slot& sl, const Scene& sc, size_t entries) {
    int64_t b = 0;
    for (const PhaseANeed& n : phase_a_needs(sl, sc, sl.batch.size(), sl.img_bytes, sl.tiles_total))
        b += fresh_bytes(n.bytes);
    const int64_t cap = fresh_bytes((int64_t)entries * 8) / 8;
    b += 2 * fresh_bytes((int64_t)entries * 8);
    b += fresh_bytes((int64_t)radix_sort_temp_bytes((int32_t)std::min<int64_t>(cap, (1 << 30) - 1), 0, sl.tile_bits));
    return b;
}

int64_t workspace_bytes(Slot& sl) {
    Workspace& w = sl.ws;
    int64_t b = 0;
    for (const Dev* d : w.devs())
        b += (int64_t)d->bytes;
    return b;
}

void release_workspace(gr_server* s, Slot& sl) {
    Workspace& w = sl.ws;
    for (Dev* d : w.devs())
        d->release(s->workspace);
    ++w.generation;
}

// Without an HBM budget the slot's phase-A buffers are sized for a full batch of the largest
// scene added so far (as many views as a batch of it may hold, up to the tile and image
// sizes a full batch can reach), so neither fuller batches nor larger scenes reallocate them
// later: each reallocation holds the scheduler for about half a millisecond.  Under a budget
// every buffer is sized for the batch at hand (exact admission).
gr_status ensure_phase_a(gr_server* s, Slot& sl, const Scene& sc) {
    Workspace& w = sl.ws;
    const size_t V = sl.batch.size();
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<PhaseANeed> needs = phase_a_needs(sl, sc, V, sl.img_bytes, sl.tiles_total);
    if (s->cfg.hbm_budget_bytes <= 0) {
        Scene largest;   // sizes only
        largest.n = std::max(sc.n, s->largest_n.load());
        const std::vector<PhaseANeed> full =
            phase_a_needs(sl, largest, (size_t)max_views_for(s, largest.n),
                          (int64_t)kBatchTiles * kTileSize * kTileSize * 4, kBatchTiles);
        for (size_t i = 0; i < needs.size(); ++i) needs[i].bytes = std::max(needs[i].bytes, full[i].bytes);
    }
    int grows = 0;
    for (const PhaseANeed& n : needs)
        S_CHECK(n.dev->ensure((size_t)n.bytes, s->workspace, &grows));
    // the camera block and the K word are copy nodes of the captured graphs: sized once (for the
    // largest batch), and a reallocation still retires every graph of the slot
    bool captured_grew = false, staging_grew = false;
    S_CHECK(w.views_host.ensure(std::max<size_t>(V, (size_t)s->cfg.max_views_per_batch) * sizeof(ViewDesc),
                                &captured_grew));
    S_CHECK(w.k_host.ensure(64, &captured_grew));
    // image staging (pinned host memory, outside the HBM budget): sized once for a full batch
    const size_t staging_full = (size_t)kBatchTiles * kTileSize * kTileSize * 4;
    S_CHECK(w.img_host[sl.img_cur].ensure(std::max<size_t>((size_t)sl.img_bytes, staging_full), &staging_grew));
    if (grows || captured_grew) ++w.generation;
    s->note_realloc(grows + (captured_grew ? 1 : 0) + (staging_grew ? 1 : 0), t0);
    s->note_peak();
    return GR_OK;
}

// Record walk of the batch (count in phase A, emit in phase B).
BuildArgs record_build_args(Slot& sl, const Scene& sc) {
    Workspace& w = sl.ws;
    const int V = (int)sl.batch.size();
    const int N = sc.n;

    BuildArgs b{};
    b.entries = w.list0.as<uint64_t>();
    b.records = w.records.as<float4>();
    b.record_slots = w.slots.as<uint32_t>();
    b.record_total = w.blocks.as<uint32_t>() + front_blocks(N);
    b.depth_keys = w.dkeys0.as<uint32_t>();
    b.depth_words = w.dorder0.as<uint32_t>();
    b.emit_records = w.emit_recs.as<uint4>();
    b.sorted_words = (sl.dsel ? w.dorder1 : w.dorder0).as<uint32_t>();
    b.depth_prefix = w.rec_prefix.as<uint32_t>();
    // the window table lives in the second list buffer, which the tile sort's first pass
    // overwrites only after the emit (capacity / kEmitWindow + 1 words in capacity * 2 words)
    b.windows = w.list1.as<uint32_t>();
    b.entry_total = w.counts.as<uint32_t>();
    b.index_bits = packed_index_bits(std::max<int64_t>((int64_t)V * N, 1));
    b.count = N;
    b.capacity = (int32_t)sl.capacity;
    b.batches = (int32_t)(((int64_t)V * N + kBuildBatch - 1) / kBuildBatch);
    b.nviews = V;
    b.v0 = w.views_host.as<ViewDesc>()[0];
    b.views = w.views.as<ViewDesc>();
    return b;
}

// Launch sequences (shared by eager runs and graph capture).
gr_status enqueue_phase_a(gr_server* s, Slot& sl, const Scene& sc) {
    Workspace& w = sl.ws;
    const int V = (int)sl.batch.size();
    const int N = sc.n;
    ViewDesc* vdev = w.views.as<ViewDesc>();
    const ViewDesc* vhost = w.views_host.as<ViewDesc>();
    S_HIP(hipMemcpyAsync(vdev, vhost, V * sizeof(ViewDesc), hipMemcpyHostToDevice, sl.stream));
    if (sc.degree == 0) {
        for (int v = 0; v < V; ++v)
            S_HIP(hipMemcpyAsync(w.color.as<float4>() + (size_t)v * N, sc.color0.p, (size_t)N * 16,
                                 hipMemcpyDeviceToDevice, sl.stream));
    }
    const int nb = front_blocks(N);
    uint32_t* vis = w.blocks.as<uint32_t>() + nb;   // records of the batch
    // the front end sets the tile ranges to -1 (tiles without entries) and zeroes the depth
    // sort's histograms, tickets and first look-back region in passing
    FillSpan dzero[2];
    uint32_t* dhist = nullptr;
    S_HIP(radix_sort_zero_spans((int32_t)std::max<int64_t>((int64_t)V * N, 1), w.dsort_temp.p, w.dsort_temp.bytes,
                                dzero, &dhist));
    FrontArgs f{};
    f.fills[0] = {w.ranges.as<uint32_t>(), 2 * (int64_t)sl.tiles_total, 0xffffffffu};
    f.fills[1] = dzero[0];
    f.fills[2] = dzero[1];
    f.nfills = 3;
    f.position = sc.position.as<float4>();
    f.scale_rotation = sc.scale_rot.as<float4>();
    f.packed = sc.packed.as<uint2>();
    f.sh = sc.sh.p;
    f.degree = sc.degree;
    f.sh_count = sc.sh_count;
    f.precision = sc.precision;
    f.color = w.color.as<float4>();
    f.clip_xy = w.xy.as<float2>();
    f.clip_z = nullptr;    // the build records carry each visible Gaussian's depth and ellipse
    f.ellipse = nullptr;
    f.conic = w.con.as<float4>();
    f.records = w.records.as<float4>();
    f.block_counts = w.blocks.as<uint32_t>();
    f.count = N;
    f.nviews = V;
    f.v0 = vhost[0];
    f.views = vdev;
    S_HIP(launch_front(f, sl.stream));
    S_HIP(launch_build_scan(w.blocks.as<uint32_t>(), nb, vis, w.blocks.as<uint32_t>() + nb + 1, sl.stream));
    S_HIP(launch_record_slots(w.blocks.as<uint32_t>(), vis, nb, kFrontThreads * V, w.slots.as<uint32_t>(),
                              sl.stream));
    BuildArgs b = record_build_args(sl, sc);
    S_HIP(launch_build_count(b, sl.stream));
    uint32_t* dk[2] = {w.dkeys0.as<uint32_t>(), w.dkeys1.as<uint32_t>()};
    int32_t* dv[2] = {w.dorder0.as<int32_t>(), w.dorder1.as<int32_t>()};
    int32_t dsel = 0;
    S_HIP(radix_sort_pairs_u32(dk, dv, &dsel, V * N, vis, 0, 32, w.dsort_temp.p, w.dsort_temp.bytes, sl.stream,
                               /*zeroed=*/true));
    sl.dsel = dsel;
    uint32_t* counts = w.counts.as<uint32_t>();
    const int vn = V * N;
    // output offsets: the counts carried by the sorted words, scanned in depth order (K); the
    // emit windows are found in phase B, once the list (which holds their table) is sized
    S_HIP(launch_build_scan_packed(reinterpret_cast<const uint32_t*>(dv[dsel]), vn, vis, w.emit_recs.as<uint4>(),
                                   b.index_bits, counts, counts + 2, w.rec_prefix.as<uint32_t>(), nullptr, 0,
                                   sl.stream));
    S_HIP(hipMemcpyAsync(w.k_host.p, counts, 4, hipMemcpyDeviceToHost, sl.stream));
    return GR_OK;
}

gr_status enqueue_phase_b(gr_server* s, Slot& sl, const Scene& sc) {
    Workspace& w = sl.ws;
    const int V = (int)sl.batch.size();
    const int N = sc.n;
    const ViewDesc* vhost = w.views_host.as<ViewDesc>();
    ViewDesc* vdev = w.views.as<ViewDesc>();
    uint32_t* counts = w.counts.as<uint32_t>();
    BuildArgs b = record_build_args(sl, sc);
    S_HIP(radix_sort_prepare((int32_t)sl.capacity, w.sort_temp.p, w.sort_temp.bytes, &b.tile_hist, sl.stream));
    b.tile_passes = (sl.tile_bits + 7) / 8;
    b.tile_digit_bits = sort_digit_bits(0, sl.tile_bits);
    if (!b.tile_passes) b.tile_hist = nullptr;   // one tile in the batch: no pass, no histograms
    S_HIP(launch_emit_windows(b.depth_prefix, b.record_total, V * N, counts, const_cast<uint32_t*>(b.windows),
                              emit_window_words(sl.capacity), sl.stream));
    S_HIP(launch_build_emit(b, sl.stream));
    uint64_t* kb[2] = {w.list0.as<uint64_t>(), w.list1.as<uint64_t>()};
    int32_t sel = 0;
    // the last pass writes the tile ranges; tiles without entries keep the -1 the front end set
    // in phase A
    RangeArgs r{};
    r.tile_shift = 0;
    r.tile_range = w.ranges.as<int32_t>();
    r.tiles = sl.tiles_total;
    S_HIP(radix_sort_keys_prepared(kb, &sel, (int32_t)sl.capacity, counts, 0, sl.tile_bits, w.sort_temp.p,
                                   w.sort_temp.bytes, sl.stream, nullptr, nullptr, &r));
    RasterArgs ra{};
    ra.tile_range = w.ranges.as<int32_t>();
    uint32_t* order = w.ranges.as<uint32_t>() + 2 * (size_t)sl.tiles_total;
    S_HIP(launch_tile_order(ra.tile_range, sl.tiles_total, order, sl.stream));
    ra.order = order;
    ra.entries = kb[sel];
    ra.clip_xy = w.xy.as<float2>();
    ra.conic = w.con.as<float4>();
    ra.color = w.color.as<float4>();
    ra.count = N;
    ra.tiles = sl.tiles_total;
    ra.nviews = V;
    ra.v0 = vhost[0];
    ra.views = vdev;
    S_HIP(launch_raster(ra, sl.stream));
    return GR_OK;
}

GraphKey graph_key(const Slot& sl, bool phase_b) {
    GraphKey k;
    k.scene = sl.scene;
    k.gen = sl.ws.generation;
    k.shape.push_back(phase_b ? 1 : 0);
    k.shape.push_back((int32_t)sl.batch.size());
    for (const Pending& p : sl.batch) {
        k.shape.push_back(p.req.width);
        k.shape.push_back(p.req.height);
    }
    if (phase_b) k.shape.push_back((int32_t)sl.capacity);
    return k;
}

void drop_graphs(Slot& sl) {
    for (auto& kv : sl.graphs) (void)hipGraphExecDestroy(kv.second);
    sl.graphs.clear();
    sl.seen.clear();
}

gr_status run_phase(gr_server* s, Slot& sl, const Scene& sc, bool phase_b) {
    auto enqueue = [&]() { return phase_b ? enqueue_phase_b(s, sl, sc) : enqueue_phase_a(s, sl, sc); };
    if (!s->cfg.use_graphs) return enqueue();
    const GraphKey key = graph_key(sl, phase_b);
    auto it = sl.graphs.find(key);
    if (it == sl.graphs.end()) {
        // a shape seen for the first time runs eagerly; its first repeat is captured for the
        // replays after it
        if (sl.seen.size() > 1024) sl.seen.clear();
        if (sl.seen[key]++ == 0) return enqueue();
        sl.seen.erase(key);
        const auto t0 = std::chrono::steady_clock::now();
        S_HIP(hipStreamBeginCapture(sl.stream, hipStreamCaptureModeThreadLocal));
        gr_status st = enqueue();
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(sl.stream, &g);
        if (st != GR_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess) return sfail(GR_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        hipGraphExec_t ex = nullptr;
        e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return sfail(GR_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        if (sl.graphs.size() > 64) drop_graphs(sl);
        it = sl.graphs.emplace(key, ex).first;
        ++s->graph_captures;
        s->capture_us +=
            (int64_t)std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
    }
    S_HIP(hipGraphLaunch(it->second, sl.stream));
    ++s->graph_launches;
    return GR_OK;
}

// Takes up to max_views pending requests of the scene with the oldest request.

bool form_batch(gr_server* s, Slot& sl) {
    int32_t best = -1;
    double oldest = 0.0;
    for (auto& kv : s->queues) {
        if (kv.second.empty()) continue;
        if (best < 0 || kv.second.front().submit_ms < oldest) {
            best = kv.first;
            oldest = kv.second.front().submit_ms;
        }
    }
    if (best < 0) return false;
    auto& q = s->queues[best];
    const Scene* sc = scene_ptr(s, best);
    int limit = sc ? max_views_for(s, sc->n) : 1;
    auto lim = s->batch_limit.find(best);
    if (lim != s->batch_limit.end()) {
        limit = std::min(limit, lim->second);
        s->batch_limit.erase(lim);
    }
    sl.batch.clear();
    // the batch's global tile ids are sorted on their bits: while they fit 16 bits the sort
    // takes two 8-bit passes, a 17th bit would add a third pass over every entry
    int64_t tiles = 0;
    while (!q.empty() && (int)sl.batch.size() < limit) {
        const FrameGeom g = make_geom(q.front().req.width, q.front().req.height);
        const int64_t t = (int64_t)g.tiles_x * g.tiles_y;
        if (!sl.batch.empty() && tiles + t > (int64_t)kBatchTiles) break;
        tiles += t;
        sl.batch.push_back(q.front());
        q.pop_front();
    }
    sl.scene = best;
    return true;
}

void requeue_front(gr_server* s, Slot& sl) {
    auto& q = s->queues[sl.scene];
    for (auto it = sl.batch.rbegin(); it != sl.batch.rend(); ++it) q.push_front(*it);
    sl.batch.clear();
    sl.state = SlotState::Idle;
}

bool other_slots_busy(gr_server* s, const Slot& me) {
    for (const Slot& o : s->slots)
        if (&o != &me && o.state != SlotState::Idle) return true;
    return false;
}

// Another slot has work on the GPU (it frees or stops claiming memory by itself).
bool other_slots_running(gr_server* s, const Slot& me) {
    for (const Slot& o : s->slots)
        if (&o != &me && (o.state == SlotState::PhaseA || o.state == SlotState::PhaseB)) return true;
    return false;
}

// Another slot, earlier in the slot list, waits for memory.
bool earlier_slot_waits(gr_server* s, const Slot& me) {
    for (const Slot& o : s->slots) {
        if (&o == &me) return false;
        if (o.state == SlotState::WaitMemory) return true;
    }
    return false;
}

void release_idle_workspaces(gr_server* s, const Slot& me) {
    for (Slot& o : s->slots) {
        if (&o == &me || o.state != SlotState::Idle) continue;
        release_workspace(s, o);
        drop_graphs(o);
    }
}

// Starts phase A of a new batch on an idle slot; returns true if it launched.
bool start_batch(gr_server* s, Slot& sl) {
    // a slot waiting for memory goes first: new batches would take the memory it waits for
    for (const Slot& o : s->slots)
        if (o.state == SlotState::WaitMemory) return false;
    // a staging buffer whose images are no longer being copied out
    std::atomic<int>* busy = &s->staging_busy[2 * (&sl - s->slots.data())];
    if (busy[sl.img_cur].load() != 0) {
        if (busy[sl.img_cur ^ 1].load() != 0) return false;
        sl.img_cur ^= 1;
    }
    if (!form_batch(s, sl)) return false;
    Scene* sc = scene_ptr(s, sl.scene);
    if (!sc) {
        gr_detail::set_last_error("scene was removed");
        complete(s, sl, GR_ERR_NOT_FOUND);
        sl.batch.clear();
        return false;
    }
    // per-view geometry, global tile numbering
    int32_t tiles = 0;
    int64_t img = 0;
    std::vector<ViewDesc> vds(sl.batch.size());
    for (size_t v = 0; v < sl.batch.size(); ++v) {
        const gr_request& r = sl.batch[v].req;
        ViewDesc& d = vds[v];
        d.cam = camera_args(r.camera);
        d.geom = make_geom(r.width, r.height);
        d.tile_base = tiles;
        d.tiles = d.geom.tiles_x * d.geom.tiles_y;
        d.out = nullptr;
        tiles += d.tiles;
        img += (int64_t)r.width * r.height * 4;
    }
    sl.tiles_total = tiles;
    sl.img_bytes = img;
    sl.tile_bits = tile_bits_for(tiles);
    std::unique_lock<std::mutex> acct(s->acct_mu);   // budget check through the allocations it admits
    const int64_t growth = phase_a_growth(sl, *sc, sl.batch.size(), img, tiles);
    if (s->over_budget(growth)) {
        if (other_slots_busy(s, sl)) {
            ++s->budget_waits;
            requeue_front(s, sl);
            return false;
        }
        release_idle_workspaces(s, sl);
        if (s->over_budget(phase_a_growth(sl, *sc, sl.batch.size(), img, tiles))) {
            // this slot's buffers are rebuilt at this batch's size (grow-only buffers may be
            // oversized from earlier, larger batches)
            release_workspace(s, sl);
            drop_graphs(sl);
        }
        while (s->over_budget(phase_a_growth(sl, *sc, sl.batch.size(), img, tiles)) && sl.batch.size() > 1) {
            // halve the batch: the tail goes back to the queue
            const size_t keep = sl.batch.size() / 2;
            auto& q = s->queues[sl.scene];
            for (size_t i = sl.batch.size(); i-- > keep;) q.push_front(sl.batch[i]);
            sl.batch.resize(keep);
            vds.resize(keep);
            tiles = 0;
            img = 0;
            for (size_t v = 0; v < keep; ++v) {
                vds[v].tile_base = tiles;
                tiles += vds[v].tiles;
                img += (int64_t)sl.batch[v].req.width * sl.batch[v].req.height * 4;
            }
            sl.tiles_total = tiles;
            sl.img_bytes = img;
            sl.tile_bits = tile_bits_for(tiles);
        }
        if (s->over_budget(phase_a_growth(sl, *sc, sl.batch.size(), img, tiles))) {
            gr_detail::set_last_error("request does not fit the HBM budget");
            sl.k = 0;
            sl.start_ms = s->now_ms();
            complete(s, sl, GR_ERR_BUDGET);
            sl.batch.clear();
            return false;
        }
    }
    if (ensure_phase_a(s, sl, *sc) != GR_OK) {
        complete(s, sl, GR_ERR_OUT_OF_MEMORY);
        sl.batch.clear();
        return false;
    }
    acct.unlock();
    int64_t off = 0;
    for (size_t v = 0; v < vds.size(); ++v) {
        vds[v].out = reinterpret_cast<uchar4*>(sl.ws.images.as<uint8_t>() + off);
        off += (int64_t)sl.batch[v].req.width * sl.batch[v].req.height * 4;
    }
    std::memcpy(sl.ws.views_host.p, vds.data(), vds.size() * sizeof(ViewDesc));
    sl.start_ms = s->now_ms();
    sl.k = 0;
    if (run_phase(s, sl, *sc, false) != GR_OK || hipEventRecord(sl.ev_a, sl.stream) != hipSuccess) {
        complete(s, sl, GR_ERR_HIP);
        sl.batch.clear();
        return false;
    }
    sl.state = SlotState::PhaseA;
    return true;
}

// Phase A finished: size the tile list and launch phase B.  Returns false if
// the slot has to wait for memory.
bool start_phase_b(gr_server* s, Slot& sl) {
    Scene* sc = scene_ptr(s, sl.scene);   // pinned by the batch's in-flight count
    Workspace& w = sl.ws;
    sl.k = w.k_host.as<uint32_t>()[0];
    if (sl.k >= (1u << 30)) {
        gr_detail::set_last_error("tile list of a batch exceeds 2^30 entries");
        complete(s, sl, GR_ERR_CAPACITY);
        sl.batch.clear();
        sl.state = SlotState::Idle;
        return true;
    }
    // bytes each tile-list buffer holds after ensure(), and the net growth for `need` entries
    auto after = [](const Dev& d, size_t want) { return d.p && d.bytes >= want ? d.bytes : std::max<size_t>(want + want / 4, 256); };
    auto growth_for = [&](size_t need) {
        const size_t cap_after = std::min(after(w.list0, need * 8), after(w.list1, need * 8)) / 8;
        const size_t sort_after =
            radix_sort_temp_bytes((int32_t)std::min<size_t>(cap_after, (1u << 30) - 1), 0, sl.tile_bits);
        return w.list0.growth(need * 8) + w.list1.growth(need * 8) + w.sort_temp.growth(sort_after);
    };
    size_t need = std::max<size_t>(sl.k, 1);
    std::unique_lock<std::mutex> acct(s->acct_mu);   // budget check through the allocations it admits
    int64_t grow = growth_for(need);
    if (grow > 0 && s->over_budget(grow)) {
        if (other_slots_running(s, sl)) {
            ++s->budget_waits;
            sl.state = SlotState::WaitMemory;
            return false;
        }
        if (earlier_slot_waits(s, sl)) {
            // nothing runs and waiting slots would wait on each other: the earliest waiting slot
            // goes on, the others hand their batches and memory back
            release_workspace(s, sl);
            drop_graphs(sl);
            s->batch_limit[sl.scene] = (int)sl.batch.size();
            requeue_front(s, sl);
            return true;
        }
        release_idle_workspaces(s, sl);
        if (s->over_budget(grow) && sl.batch.size() > 1) {
            // run again with half as many views: phase A is redone for them (the entries of the
            // batch's views interleave in depth order, so no prefix of this list is theirs)
            release_workspace(s, sl);
            drop_graphs(sl);
            s->batch_limit[sl.scene] = (int)sl.batch.size() / 2;
            requeue_front(s, sl);
            return true;
        }
        if (s->over_budget(grow) && !s->over_budget(fresh_workspace_bytes(sl, *sc, need) - workspace_bytes(sl))) {
            // phase-A buffers oversized by earlier, larger batches are what is in the way: run the
            // batch again from phase A on buffers of its own size
            release_workspace(s, sl);
            drop_graphs(sl);
            s->batch_limit[sl.scene] = (int)sl.batch.size();   // the same views again, no more
            requeue_front(s, sl);
            return true;
        }
        if (s->over_budget(grow)) {
            gr_detail::set_last_error("tile list does not fit the HBM budget");
            complete(s, sl, GR_ERR_BUDGET);
            sl.batch.clear();
            sl.state = SlotState::Idle;
            return true;
        }
    }
    const auto t0 = std::chrono::steady_clock::now();
    int grows = 0;
    if (w.list0.ensure(need * 8, s->workspace, &grows) != GR_OK ||
        w.list1.ensure(need * 8, s->workspace, &grows) != GR_OK) {
        complete(s, sl, GR_ERR_OUT_OF_MEMORY);
        sl.batch.clear();
        sl.state = SlotState::Idle;
        return true;
    }
    sl.capacity = (uint32_t)(std::min(w.list0.bytes, w.list1.bytes) / 8);
    sl.capacity = std::min<uint32_t>(sl.capacity, (1u << 30) - 1);
    const size_t sort_bytes = radix_sort_temp_bytes((int32_t)sl.capacity, 0, sl.tile_bits);
    if (w.sort_temp.ensure(sort_bytes, s->workspace, &grows) != GR_OK) {
        complete(s, sl, GR_ERR_OUT_OF_MEMORY);
        sl.batch.clear();
        sl.state = SlotState::Idle;
        return true;
    }
    acct.unlock();
    s->note_peak();
    s->note_realloc(grows, t0);
    if (grows) ++w.generation;
    // the image copy stays out of the graph: the staging buffer alternates between batches
    if (run_phase(s, sl, *sc, true) != GR_OK ||
        hipMemcpyAsync(w.img_host[sl.img_cur].p, w.images.p, (size_t)sl.img_bytes, hipMemcpyDeviceToHost,
                       sl.stream) != hipSuccess ||
        hipEventRecord(sl.ev_b, sl.stream) != hipSuccess) {
        complete(s, sl, GR_ERR_HIP);
        sl.batch.clear();
        sl.state = SlotState::Idle;
        return true;
    }
    sl.state = SlotState::PhaseB;
    return true;
}

// Hands the batch's images to the copy-out threads (views without an output buffer complete
// here) and frees the slot; the slot's other staging buffer serves its next batch.
void finish_batch(gr_server* s, Slot& sl) {
    std::atomic<int>* busy = &s->staging_busy[2 * (&sl - s->slots.data()) + sl.img_cur];
    const uint8_t* staging = sl.ws.img_host[sl.img_cur].as<uint8_t>();
    std::vector<CopyChunk> chunks;
    std::vector<gr_completion> now;
    int64_t off = 0;
    for (const Pending& p : sl.batch) {
        const size_t bytes = (size_t)p.req.width * p.req.height * 4;
        gr_completion c{};
        c.request_id = p.id;
        c.user_tag = p.req.user_tag;
        c.status = GR_OK;
        c.batch_views = (int32_t)sl.batch.size();
        c.submit_ms = p.submit_ms;
        c.start_ms = sl.start_ms;
        c.tile_list_size = sl.k;
        if (!p.req.out_rgba8) {
            c.finish_ms = s->now_ms();
            now.push_back(c);
        } else {
            auto v = std::make_shared<ViewCopy>();
            v->c = c;
            v->scene = sl.scene;
            v->staging_busy = busy;
            const size_t n = (bytes + kCopyChunk - 1) / kCopyChunk;
            v->left.store((int)n);
            busy->fetch_add(1);
            for (size_t k = 0; k < n; ++k) {
                const size_t at = k * kCopyChunk;
                chunks.push_back({staging + off + at, p.req.out_rgba8 + at, std::min(kCopyChunk, bytes - at), v});
            }
        }
        off += (int64_t)bytes;
    }
    ++s->batches;
    s->views += (int64_t)sl.batch.size();
    if (!now.empty()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (const gr_completion& c : now) s->done.push_back(c);
            auto it = s->scenes.find(sl.scene);
            if (it != s->scenes.end()) it->second->inflight -= (int)now.size();
        }
        s->cv_done.notify_all();
    }
    if (!chunks.empty()) {
        {
            std::lock_guard<std::mutex> lk(s->copy_mu);
            for (CopyChunk& ch : chunks) s->copy_q.push_back(std::move(ch));
        }
        s->copy_cv.notify_all();
    }
    sl.batch.clear();
    sl.state = SlotState::Idle;
    sl.img_cur ^= 1;
}

void scheduler(gr_server* s) {
    (void)hipSetDevice(s->device);
    while (!s->stop.load()) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            bool busy = false;
            for (const Slot& sl : s->slots) busy |= sl.state != SlotState::Idle;
            bool queued = false;
            for (auto& kv : s->queues) queued |= !kv.second.empty();
            if (!busy && !queued && s->incoming.empty())
                s->cv_submit.wait_for(lk, std::chrono::milliseconds(2));
            while (!s->incoming.empty()) {
                Pending p = s->incoming.front();
                s->incoming.pop_front();
                s->queues[p.req.scene].push_back(p);
            }
        }
        bool progress = false;
        for (Slot& sl : s->slots) {
            if (sl.state == SlotState::PhaseA && hipEventQuery(sl.ev_a) == hipSuccess) {
                progress |= start_phase_b(s, sl);
            } else if (sl.state == SlotState::WaitMemory) {
                progress |= start_phase_b(s, sl);
            } else if (sl.state == SlotState::PhaseB && hipEventQuery(sl.ev_b) == hipSuccess) {
                finish_batch(s, sl);
                progress = true;
            }
        }
        for (Slot& sl : s->slots)
            if (sl.state == SlotState::Idle) progress |= start_batch(s, sl);
        if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

}  // namespace

extern "C" {

gr_status gr_server_create(gr_server** out, const gr_server_config* cfg) {
    if (!out || !cfg || cfg->slots < 1 || cfg->slots > 16 || cfg->max_views_per_batch < 1 ||
        cfg->max_views_per_batch > 1024 || cfg->hbm_budget_bytes < 0)
        return sfail(GR_ERR_INVALID_ARGUMENT, "gr_server_create: bad configuration");
    if (cfg->device >= 0) S_HIP(hipSetDevice(cfg->device));
    auto* s = new gr_server();
    s->cfg = *cfg;
    S_HIP(hipGetDevice(&s->device));
    s->t0 = std::chrono::steady_clock::now();
    s->slots.resize(cfg->slots);
    s->staging_busy.reset(new std::atomic<int>[2 * cfg->slots]);
    for (int i = 0; i < 2 * cfg->slots; ++i) s->staging_busy[i].store(0);
    for (Slot& sl : s->slots) {
        S_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
        S_HIP(hipEventCreateWithFlags(&sl.ev_a, hipEventDisableTiming));
        S_HIP(hipEventCreateWithFlags(&sl.ev_b, hipEventDisableTiming));
    }
    for (int i = 0; i < kCopyThreads; ++i) s->copiers.emplace_back(copier, s);
    s->worker = std::thread(scheduler, s);
    *out = s;
    return GR_OK;
}

gr_status gr_server_destroy(gr_server* s) {
    if (!s) return GR_OK;
    s->stop = true;
    s->cv_submit.notify_all();
    if (s->worker.joinable()) s->worker.join();
    {
        std::lock_guard<std::mutex> lk(s->copy_mu);
        s->copy_stop = true;
    }
    s->copy_cv.notify_all();
    for (std::thread& t : s->copiers) t.join();   // drains the copy queue first
    (void)hipSetDevice(s->device);
    for (Slot& sl : s->slots) {
        (void)hipStreamSynchronize(sl.stream);
        drop_graphs(sl);
        Workspace& w = sl.ws;
        for (Dev* d : w.devs())
            d->release(s->workspace);
        w.views_host.release();
        w.k_host.release();
        w.img_host[0].release();
        w.img_host[1].release();
        (void)hipEventDestroy(sl.ev_a);
        (void)hipEventDestroy(sl.ev_b);
        (void)hipStreamDestroy(sl.stream);
    }
    for (auto& kv : s->scenes) {
        kv.second->release(s->resident);
        delete kv.second;
    }
    delete s;
    return GR_OK;
}

gr_status gr_server_add_scene(gr_server* s, const gr_model_host* m, int32_t* scene_id) {
    if (!s || !m || !scene_id || m->num_gaussians <= 0 || m->sh_degree < 0 || m->sh_degree > 3 ||
        m->precision < 0 || m->precision > 2 || !m->position || !m->scale_and_rotation ||
        (m->sh_degree > 0 && !m->sh) || (m->sh_degree == 0 && !m->color))
        return sfail(GR_ERR_INVALID_ARGUMENT, "gr_server_add_scene: bad model");
    (void)hipSetDevice(s->device);
    auto sc = std::make_unique<Scene>();
    const int n = m->num_gaussians;
    sc->n = n;
    sc->degree = m->sh_degree;
    sc->sh_count = 3 * (m->sh_degree + 1) * (m->sh_degree + 1);
    sc->precision = m->precision;
    const size_t groups = ((size_t)n + kShGroup - 1) / kShGroup;
    const size_t shn = groups * kShGroup * (size_t)sc->sh_count;
    const size_t pos_b = (size_t)n * 16;
    const size_t sr_b = m->precision ? 0 : (size_t)n * 16;
    const size_t pk_b = m->precision ? (size_t)n * 8 : 0;
    const size_t sh_b = m->sh_degree > 0 ? shn * (m->precision ? 2 : 4) : 0;
    const size_t c0_b = m->sh_degree == 0 ? (size_t)n * 16 : 0;
    const int64_t bytes = (int64_t)(pos_b + sr_b + pk_b + sh_b + c0_b);
    std::atomic<int64_t>& acct = s->resident;
    // frees whatever was allocated when a later step fails (and un-counts it)
    struct Guard {
        Scene* sc;
        std::atomic<int64_t>& acct;
        ~Guard() {
            if (sc) sc->release(acct);
        }
    } guard{sc.get(), acct};
    {
        // the buffers take exactly the admitted bytes; the check and the allocations run under
        // the admission lock, so a concurrent workspace growth cannot use the same headroom
        std::lock_guard<std::mutex> alk(s->acct_mu);
        if (s->over_budget(bytes)) return sfail(GR_ERR_BUDGET, "scene does not fit the HBM budget");
        S_CHECK(sc->position.ensure(pos_b, acct, nullptr, true));
        if (sr_b) S_CHECK(sc->scale_rot.ensure(sr_b, acct, nullptr, true));
        if (pk_b) S_CHECK(sc->packed.ensure(pk_b, acct, nullptr, true));
        if (sh_b) S_CHECK(sc->sh.ensure(sh_b, acct, nullptr, true));
        if (c0_b) S_CHECK(sc->color0.ensure(c0_b, acct, nullptr, true));
    }
    if (m->precision) {
        std::vector<float> pos((size_t)4 * n);
        std::vector<uint16_t> pk((size_t)4 * n);
        std::vector<uint16_t> sh16(m->sh_degree > 0 ? shn : 1);
        S_CHECK(gr_pack_precision_model(m, m->precision, pos.data(), pk.data(), sh16.data()));
        S_HIP(hipMemcpy(sc->position.p, pos.data(), pos_b, hipMemcpyHostToDevice));
        S_HIP(hipMemcpy(sc->packed.p, pk.data(), pk_b, hipMemcpyHostToDevice));
        if (sh_b) S_HIP(hipMemcpy(sc->sh.p, sh16.data(), sh_b, hipMemcpyHostToDevice));
    } else {
        S_HIP(hipMemcpy(sc->position.p, m->position, pos_b, hipMemcpyHostToDevice));
        S_HIP(hipMemcpy(sc->scale_rot.p, m->scale_and_rotation, sr_b, hipMemcpyHostToDevice));
        if (sh_b) {
            std::vector<float> al(shn);
            S_CHECK(gr_realign_spherical_harmonics(m->sh, al.data(), kShGroup, sc->sh_count, n));
            S_HIP(hipMemcpy(sc->sh.p, al.data(), sh_b, hipMemcpyHostToDevice));
        }
    }
    if (c0_b) S_HIP(hipMemcpy(sc->color0.p, m->color, c0_b, hipMemcpyHostToDevice));
    guard.sc = nullptr;
    s->note_peak();
    std::lock_guard<std::mutex> lk(s->mu);
    sc->id = s->next_scene++;
    sc->bytes = bytes;
    *scene_id = sc->id;
    const int32_t id = sc->id;   // (C++17 sequences the right operand of = first)
    if (n > s->largest_n.load()) s->largest_n.store(n);   // writers hold mu
    s->scenes[id] = sc.release();
    return GR_OK;
}

gr_status gr_server_remove_scene(gr_server* s, int32_t id) {
    if (!s) return sfail(GR_ERR_INVALID_ARGUMENT, "server is NULL");
    Scene* sc = nullptr;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->scenes.find(id);
        if (it == s->scenes.end()) return sfail(GR_ERR_NOT_FOUND, "no such scene");
        sc = it->second;
        sc->removing = true;
        s->cv_done.wait(lk, [&] { return sc->inflight == 0; });
        s->scenes.erase(it);
    }
    // graphs are keyed by scene id and ids are never reused, so stale graphs are never replayed
    sc->release(s->resident);
    delete sc;
    return GR_OK;
}

gr_status gr_server_submit(gr_server* s, const gr_request* r, uint64_t* request_id) {
    if (!s || !r || r->width <= 0 || r->height <= 0 || r->width > 16384 || r->height > 16384)
        return sfail(GR_ERR_INVALID_ARGUMENT, "gr_server_submit: bad request");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->scenes.find(r->scene);
    if (it == s->scenes.end() || it->second->removing) return sfail(GR_ERR_NOT_FOUND, "no such scene");
    const FrameGeom g = make_geom(r->width, r->height);
    if ((int64_t)g.tiles_x * g.tiles_y * s->cfg.max_views_per_batch >= (1 << 24))
        return sfail(GR_ERR_INVALID_ARGUMENT, "frame too large for the batch tile space");
    Pending p;
    p.id = s->next_request++;
    p.req = *r;
    p.submit_ms = s->now_ms();
    it->second->inflight += 1;
    s->incoming.push_back(p);
    if (request_id) *request_id = p.id;
    s->cv_submit.notify_one();
    return GR_OK;
}

gr_status gr_server_poll(gr_server* s, gr_completion* out, int32_t max, int32_t* n_out, int32_t wait_ms) {
    if (!s || !n_out || (max > 0 && !out)) return sfail(GR_ERR_INVALID_ARGUMENT, "gr_server_poll: bad argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty() && wait_ms > 0)
        s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !s->done.empty(); });
    int32_t n = 0;
    while (n < max && !s->done.empty()) {
        out[n++] = s->done.front();
        s->done.pop_front();
    }
    *n_out = n;
    return GR_OK;
}

gr_status gr_server_get_stats(gr_server* s, gr_server_stats* o) {
    if (!s || !o) return sfail(GR_ERR_INVALID_ARGUMENT, "gr_server_get_stats: bad argument");
    o->resident_bytes = s->resident.load();
    o->workspace_bytes = s->workspace.load();
    o->peak_bytes = s->peak.load();
    o->batches = s->batches.load();
    o->views = s->views.load();
    o->graph_launches = s->graph_launches.load();
    o->budget_waits = s->budget_waits.load();
    o->graph_captures = s->graph_captures.load();
    o->capture_ms = (double)s->capture_us.load() / 1000.0;
    o->copy_out_ms = (double)s->copy_us.load() / 1000.0;
    o->workspace_reallocs = s->reallocs.load();
    o->realloc_ms = (double)s->realloc_us.load() / 1000.0;
    return GR_OK;
}

double gr_server_clock_ms(gr_server* s) { return s ? s->now_ms() : 0.0; }

gr_status gr_server_last_error(gr_server* s, char* buf, int32_t buf_bytes) {
    if (!s || !buf || buf_bytes <= 0) return sfail(GR_ERR_INVALID_ARGUMENT, "gr_server_last_error: bad argument");
    std::lock_guard<std::mutex> lk(s->err_mu);
    const size_t n = std::min<size_t>(s->last_error.size(), (size_t)buf_bytes - 1);
    std::memcpy(buf, s->last_error.data(), n);
    buf[n] = '\0';
    return GR_OK;
}

}  // extern "C"
