// mock_hip.cpp -- host-only stand-in for the HIP runtime and for the stage
// launchers, so that the serving scheduler (gr_server.hip, compiled with
// --cuda-host-only) runs on a machine without a GPU, under AddressSanitizer.
//
// Test infrastructure only.  "Device" memory is host memory; every mocked
// kernel reads or writes exactly the byte ranges the real kernel would, so an
// undersized workspace, a stale pointer replayed from a graph or a buffer freed
// while in use shows up as an ASAN report.  Graph capture records the enqueued
// operations as closures and replays them on hipGraphLaunch, the way a real
// graph replays its captured nodes (pointers and kernel arguments frozen at
// capture time, memcpy sources re-read at replay).
//
// Mocked raster: every pixel of view v is set to (uint8)cam.position.x of that
// view, so the test can check that each request received its own image.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "gr_capi.h"
#include "gr_kernels.hpp"

namespace {

struct MockStream {
    bool capturing = false;
    std::vector<std::function<void()>>* rec = nullptr;
};

struct MockGraph {
    std::vector<std::function<void()>> ops;
};

void run_or_record(hipStream_t s, std::function<void()> f) {
    MockStream* ms = reinterpret_cast<MockStream*>(s);
    if (ms && ms->capturing) ms->rec->push_back(std::move(f));
    else f();
}

void touch(void* p, size_t n) {
    if (n) std::memset(p, 0, n);
}

void rd(const void* p, size_t n) {
    if (n) {
        volatile const char* c = static_cast<const char*>(p);
        (void)c[0];
        (void)c[n - 1];
    }
}

std::mutex g_err_mu;
std::string g_last_error;

}  // namespace

namespace gr_detail {
void set_last_error(const std::string& msg) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_last_error = msg;
}
std::string get_last_error() {
    std::lock_guard<std::mutex> lk(g_err_mu);
    return g_last_error;
}
}  // namespace gr_detail

extern "C" {

hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDevice(int* d) {
    *d = 0;
    return hipSuccess;
}
const char* hipGetErrorString(hipError_t) { return "mock error"; }
// Allocation failure injection: after mock_fail_malloc_after(k), the next k device
// allocations succeed and the ones after them fail with hipErrorOutOfMemory (k < 0: none fail).
static std::mutex g_fail_mu;
static int g_malloc_budget = -1;
void mock_fail_malloc_after(int k) {
    std::lock_guard<std::mutex> lk(g_fail_mu);
    g_malloc_budget = k;
}
hipError_t hipMalloc(void** p, size_t n) {
    {
        std::lock_guard<std::mutex> lk(g_fail_mu);
        if (g_malloc_budget == 0) {
            *p = nullptr;
            return hipErrorOutOfMemory;
        }
        if (g_malloc_budget > 0) --g_malloc_budget;
    }
    *p = std::malloc(n);
    std::memset(*p, 0xcd, n);
    return hipSuccess;
}
hipError_t hipFree(void* p) {
    std::free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t n, unsigned int) {
    *p = std::malloc(n);
    std::memset(*p, 0xcd, n);
    return hipSuccess;
}
hipError_t hipHostFree(void* p) {
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
    std::memcpy(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t st) {
    run_or_record(st, [=] { std::memcpy(d, s, n); });
    return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
    *s = reinterpret_cast<hipStream_t>(new MockStream());
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
    delete reinterpret_cast<MockStream*>(s);
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned int) {
    *e = reinterpret_cast<hipEvent_t>(new int(0));
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
    delete reinterpret_cast<int*>(e);
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
    MockStream* ms = reinterpret_cast<MockStream*>(s);
    if (ms->capturing) return hipErrorInvalidValue;
    ms->capturing = true;
    ms->rec = new std::vector<std::function<void()>>();
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
    MockStream* ms = reinterpret_cast<MockStream*>(s);
    if (!ms->capturing) return hipErrorInvalidValue;
    MockGraph* mg = new MockGraph();
    mg->ops = std::move(*ms->rec);
    delete ms->rec;
    ms->rec = nullptr;
    ms->capturing = false;
    *g = reinterpret_cast<hipGraph_t>(mg);
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* ex, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
    *ex = reinterpret_cast<hipGraphExec_t>(new MockGraph(*reinterpret_cast<MockGraph*>(g)));
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
    delete reinterpret_cast<MockGraph*>(g);
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t g) {
    delete reinterpret_cast<MockGraph*>(g);
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t g, hipStream_t) {
    for (auto& f : reinterpret_cast<MockGraph*>(g)->ops) f();
    return hipSuccess;
}

// host helpers the server calls (the real ones live in gr_api.hip)
gr_status gr_realign_spherical_harmonics(const float* src, float* dst, int32_t group_size, int32_t sh_count,
                                         int32_t n) {
    const int groups = (n + group_size - 1) / group_size;
    for (int grp = 0; grp < groups; ++grp) {
        const size_t start = (size_t)grp * group_size * sh_count;
        const int here = std::min(group_size, n - grp * group_size);
        for (int i = 0; i < sh_count; ++i) {
            float* d = dst + start + (size_t)group_size * i;
            for (int k = 0; k < here; ++k) d[k] = src[start + (size_t)sh_count * k + i];
            for (int k = here; k < group_size; ++k) d[k] = 0.0f;
        }
    }
    return GR_OK;
}

gr_status gr_pack_precision_model(const gr_model_host* m, int32_t, float* pos, uint16_t* packed, uint16_t* sh) {
    const size_t n = (size_t)m->num_gaussians;
    std::memset(pos, 0, n * 16);
    std::memset(packed, 0, n * 8);
    if (m->sh_degree > 0) {
        const size_t groups = (n + 255) / 256;
        std::memset(sh, 0, groups * 256 * 3 * (m->sh_degree + 1) * (m->sh_degree + 1) * 2);
    }
    return GR_OK;
}

const char* mock_last_error() { return g_last_error.c_str(); }

}  // extern "C"

namespace gr {

hipError_t launch_sh(const ShArgs& a, int, int precision, hipStream_t s) {
    run_or_record(s, [=] {
        rd(a.position, (size_t)a.count * 16);
        if (precision) rd(a.packed, (size_t)a.count * 8);
        for (int v = 0; v < a.nviews; ++v) rd(&view_of(v, a.v0, a.views), sizeof(ViewDesc));
        touch(a.color, (size_t)a.nviews * a.count * 16);
    });
    return hipSuccess;
}

hipError_t launch_project(const ProjectArgs& a, hipStream_t s) {
    run_or_record(s, [=] {
        const size_t n = (size_t)a.nviews * a.count;
        rd(a.position, (size_t)a.count * 16);
        touch(a.clip_xy, n * 8);
        touch(a.clip_z, n * 4);
        touch(a.ellipse, n * 16);
        touch(a.conic, n * 16);
    });
    return hipSuccess;
}

// Front end: every Gaussian of every view visible; each block's records in (view, index) order
// in its own staging slice.
hipError_t launch_front(const FrontArgs& a, hipStream_t s) {
    if (a.nviews < 1 || a.nviews > kFrontMaxViews || !a.records || !a.block_counts || !a.clip_xy || !a.conic ||
        (a.degree > 0 && (!a.sh || !a.color)))
        return hipErrorInvalidValue;
    if (a.nfills < 0 || a.nfills > kFrontFills) return hipErrorInvalidValue;
    run_or_record(s, [=] {
        for (int f = 0; f < a.nfills; ++f)
            for (int64_t j = 0; j < a.fills[f].n; ++j) a.fills[f].p[j] = a.fills[f].value;
        const size_t n = (size_t)a.count, vn = (size_t)a.nviews * n;
        rd(a.position, n * 16);
        if (a.precision) rd(a.packed, n * 8); else rd(a.scale_rotation, n * 16);
        if (a.degree > 0) {
            rd(a.sh, ((n + 255) / 256) * 256 * (size_t)a.sh_count * (a.precision ? 2 : 4));
            touch(a.color, vn * 16);
        }
        touch(a.clip_xy, vn * 8);
        if (a.clip_z) touch(a.clip_z, vn * 4);
        if (a.ellipse) touch(a.ellipse, vn * 16);
        touch(a.conic, vn * 16);
        const int nb = front_blocks(a.count);
        for (int blk = 0; blk < nb; ++blk) {
            const uint32_t here = (uint32_t)std::min(kFrontThreads, a.count - kFrontThreads * blk);
            const uint32_t c = here * (uint32_t)a.nviews;
            touch(a.records + 2 * (size_t)blk * kFrontThreads * a.nviews, (size_t)c * 32);
            a.block_counts[blk] = c;
        }
    });
    return hipSuccess;
}

hipError_t launch_record_slots(const uint32_t* offsets, const uint32_t* total, int32_t blocks, int32_t region,
                               uint32_t* slots, hipStream_t s) {
    if (blocks <= 0) return hipSuccess;
    if (!offsets || !total || !slots || region <= 0 || (int64_t)blocks * region > (int64_t)UINT32_MAX)
        return hipErrorInvalidValue;
    run_or_record(s, [=] {
        for (int b = 0; b < blocks; ++b) {
            const uint32_t o = offsets[b], e = b + 1 < blocks ? offsets[b + 1] : *total;
            if (e < o || e - o > (uint32_t)region) std::abort();   // a block holds at most `region` records
            for (uint32_t r = 0; o + r < e; ++r) slots[o + r] = (uint32_t)b * (uint32_t)region + r;
        }
    });
    return hipSuccess;
}

// index-order walk: 37 entries per build batch of every view; record walk: 3 entries per record
hipError_t launch_build_count(const BuildArgs& a, hipStream_t s) {
    run_or_record(s, [=] {
        if (a.records) {
            const uint32_t nvis = std::min<uint32_t>(*a.record_total, (uint32_t)(a.nviews * a.count));
            if (a.index_bits < 1 || a.index_bits > 31 || ((int64_t)1 << a.index_bits) < (int64_t)a.nviews * a.count)
                std::abort();
            for (uint32_t q = 0; q < nvis; ++q) {
                rd(&a.records[2 * (size_t)a.record_slots[q]], 32);
                // depth word 7, 3 entries: 4 candidates in a 4-wide rect at the corner, 3 of them pass
                a.depth_keys[q] = 7u;
                a.depth_words[q] = q | (3u << a.index_bits);
                a.emit_records[q] = make_uint4(0xbu, 0u, q, 3u << kEmitCornerBits);
            }
            return;
        }
        const size_t n = (size_t)a.nviews * a.count;
        rd(a.clip_xy, n * 8);
        rd(a.clip_z, n * 4);
        rd(a.ellipse, n * 16);
        for (int i = 0; i < a.nviews * a.batches; ++i) a.batch_counts[i] = 37;
    });
    return hipSuccess;
}

int packed_index_bits(int64_t n) {
    int b = 1;
    while (b < 31 && ((int64_t)1 << b) < n) ++b;
    return b;
}

// windows[w] = the position whose entries [prefix[q], prefix[q + 1]) hold entry w * kEmitWindow
static void mock_windows(const uint32_t* prefix, uint32_t nv, uint32_t total, uint32_t* windows, int64_t nwin) {
    for (uint32_t q = 0; q < nv; ++q) {
        const uint64_t end = q + 1 < nv ? prefix[q + 1] : total;
        const uint64_t ew = (uint64_t)kEmitWindow;
        for (uint64_t b = ((uint64_t)prefix[q] + ew - 1) / ew; b * ew < end && (int64_t)b < nwin; ++b)
            windows[b] = q;
    }
}

hipError_t launch_build_scan_packed(const uint32_t* words, int32_t n, const uint32_t* valid, const uint4* erec,
                                    int32_t index_bits, uint32_t* total, uint32_t* scratch, uint32_t* prefix,
                                    uint32_t* windows, int64_t nwindows, hipStream_t s) {
    if (index_bits < 1 || index_bits > 31 || (n > 0 && ((int64_t)1 << index_bits) < n)) return hipErrorInvalidValue;
    if (nwindows < 0 || (nwindows > 0 && !windows)) return hipErrorInvalidValue;
    run_or_record(s, [=] {
        const uint32_t nv = std::min<uint32_t>(*valid, (uint32_t)std::max(n, 0));
        const uint32_t esc = 0xffffffffu >> index_bits;
        uint32_t acc = 0;
        for (uint32_t i = 0; i < nv; ++i) {
            const uint32_t w = words[i];
            const uint32_t idx = w & ((1u << index_bits) - 1u);
            if (idx >= nv) std::abort();   // a pair names a record that does not exist
            const uint32_t c = (w >> index_bits) == esc ? (uint32_t)__builtin_popcount(erec[idx].x) : (w >> index_bits);
            prefix[i] = acc;
            acc += c;
        }
        *total = acc;
        if (nwindows > 0) mock_windows(prefix, nv, acc, windows, nwindows);
        touch(scratch, (size_t)scan_scratch_words(n) * 4);
    });
    return hipSuccess;
}

hipError_t launch_emit_windows(const uint32_t* prefix, const uint32_t* valid, int32_t n, const uint32_t* total,
                               uint32_t* windows, int64_t nwindows, hipStream_t s) {
    if (n <= 0 || nwindows <= 0) return hipSuccess;
    if (!prefix || !valid || !total || !windows) return hipErrorInvalidValue;
    run_or_record(s, [=] {
        const uint32_t nv = std::min<uint32_t>(*valid, (uint32_t)n);
        rd(prefix, (size_t)nv * 4);
        touch(windows, (size_t)std::min<int64_t>(nwindows, (int64_t)*total / (uint32_t)kEmitWindow + 1) * 4);
        mock_windows(prefix, nv, *total, windows, nwindows);
    });
    return hipSuccess;
}

int scan_scratch_words(int32_t n) { return (n + 1023) / 1024 + 64; }

hipError_t launch_build_scan(uint32_t* data, int32_t n, uint32_t* total, uint32_t* scratch, hipStream_t s) {
    run_or_record(s, [=] {
        uint32_t acc = 0;
        for (int i = 0; i < n; ++i) {
            const uint32_t c = data[i];
            data[i] = acc;
            acc += c;
        }
        *total = acc;
        touch(scratch, (size_t)scan_scratch_words(n) * 4);
    });
    return hipSuccess;
}

hipError_t launch_build_emit(const BuildArgs& a, hipStream_t s) {
    run_or_record(s, [=] {
        uint32_t K;
        if (a.records) {
            const uint32_t nvis = std::min<uint32_t>(*a.record_total, (uint32_t)(a.nviews * a.count));
            K = 0;
            for (uint32_t q = 0; q < nvis; ++q) {
                const uint32_t rec = a.sorted_words[q] & ((1u << a.index_bits) - 1u);
                if (rec >= nvis) std::abort();
                const uint4 e = a.emit_records[rec];
                if (a.depth_prefix[q] > 3u * nvis) std::abort();   // offsets lie inside the list
                if ((e.w & kEmitBig) || e.y != 0u || e.z != rec) std::abort();
                K += (uint32_t)__builtin_popcount(e.x);
            }
            if (K != *a.entry_total) std::abort();
            // every window's first record holds the window's first entry
            for (uint32_t w0 = 0; w0 * (uint32_t)kEmitWindow < std::min<uint32_t>(K, (uint32_t)a.capacity); ++w0) {
                const uint32_t q = a.windows[w0];
                if (q >= nvis || a.depth_prefix[q] > w0 * (uint32_t)kEmitWindow) std::abort();
            }
            if (a.tile_hist)
                for (int k = 0; k < a.tile_passes; ++k) a.tile_hist[k * 256] += std::min<uint32_t>(K, (uint32_t)a.capacity);
        } else {
            K = a.batch_counts[a.nviews * a.batches];
        }
        const uint32_t n = std::min<uint32_t>(K, (uint32_t)a.capacity);
        if (a.records) {
            touch(a.entries, (size_t)n * 8);
        } else {
            touch(a.keys, (size_t)n * 8);
            touch(a.values, (size_t)n * 4);
        }
    });
    return hipSuccess;
}

hipError_t launch_fill_i32(int32_t* p, int32_t value, int64_t n, hipStream_t s) {
    run_or_record(s, [=] {
        for (int64_t i = 0; i < n; ++i) p[i] = value;
    });
    return hipSuccess;
}

constexpr int32_t kNoRanges = -7;   // marks a list that did not fit (the real ranges stay empty)

hipError_t launch_ranges(const RangeArgs& a, hipStream_t s) {
    run_or_record(s, [=] {
        if (a.tile_shift != 0 && a.tile_shift != 32) std::abort();
        if (a.size_dev && *a.size_dev > (uint32_t)a.size) {   // K beyond the list: nothing is ranged
            a.tile_range[0] = kNoRanges;
            return;
        }
        const uint32_t n = a.size_dev ? *a.size_dev : (uint32_t)a.size;
        rd(a.keys, (size_t)n * 8);
        touch(a.tile_range, (size_t)a.tiles * 8);
    });
    return hipSuccess;
}

hipError_t launch_tile_order(const int32_t* tile_range, int32_t tiles, uint32_t* order, hipStream_t s) {
    if (tiles <= 0) return hipSuccess;
    if (!tile_range || !order) return hipErrorInvalidValue;
    run_or_record(s, [=] {
        rd(tile_range, (size_t)tiles * 8);
        for (int32_t t = 0; t < tiles; ++t) order[t] = (uint32_t)t;
    });
    return hipSuccess;
}

hipError_t launch_raster(const RasterArgs& a, hipStream_t s) {
    run_or_record(s, [=] {
        rd(a.tile_range, (size_t)a.tiles * 8);
        if (a.order) {   // a permutation of the tiles
            std::vector<char> seen((size_t)a.tiles, 0);
            for (int32_t t = 0; t < a.tiles; ++t) {
                if (a.order[t] >= (uint32_t)a.tiles || seen[a.order[t]]) std::abort();
                seen[a.order[t]] = 1;
            }
        }
        const bool empty = a.tile_range[0] == kNoRanges;   // every tile empty: RGBA 0
        int tiles = 0;
        for (int v = 0; v < a.nviews; ++v) {
            const ViewDesc& d = view_of(v, a.v0, a.views);
            if (d.tile_base != tiles) std::abort();   // views must tile the global tile space in order
            tiles += d.tiles;
            std::memset(d.out, empty ? 0 : (int)(uint8_t)d.cam.position.x, (size_t)d.geom.width * d.geom.height * 4);
        }
        if (tiles != a.tiles) std::abort();
    });
    return hipSuccess;
}

template <typename KeyT>
hipError_t mock_sort(KeyT* keys[2], int32_t* values[2], int32_t* sel, int32_t count, const uint32_t* count_dev,
                     int32_t begin_bit, int32_t end_bit, void* temp, size_t temp_bytes, hipStream_t s);

hipError_t radix_sort_prepare(int32_t count, void* temp, size_t temp_bytes, uint32_t** hist, hipStream_t s) {
    if (count < 1) count = 1;
    if (temp_bytes < (size_t)count * 4 + 4096 || !temp) return hipErrorInvalidValue;
    *hist = static_cast<uint32_t*>(temp);
    const size_t zero = std::min<size_t>(temp_bytes, 8 * 256 * 4);
    run_or_record(s, [=] { std::memset(temp, 0, zero); });
    return hipSuccess;
}

int sort_digit_bits(int32_t begin_bit, int32_t end_bit) {
    const int p = (end_bit - begin_bit + 7) / 8;
    return p > 0 ? (end_bit - begin_bit + p - 1) / p : 0;
}

size_t radix_sort_temp_bytes(int32_t count, int32_t, int32_t end_bit) {
    return (size_t)count * 4 + 4096 + (size_t)end_bit;
}

hipError_t radix_sort_pairs(uint64_t* keys[2], int32_t* values[2], int32_t* sel, int32_t count,
                            const uint32_t* count_dev, int32_t begin_bit, int32_t end_bit, void* temp,
                            size_t temp_bytes, hipStream_t s) {
    return mock_sort(keys, values, sel, count, count_dev, begin_bit, end_bit, temp, temp_bytes, s);
}

hipError_t radix_sort_keys_prepared(uint64_t* keys[2], int32_t* sel, int32_t count, const uint32_t* count_dev,
                                    int32_t begin_bit, int32_t end_bit, void* temp, size_t temp_bytes, hipStream_t s,
                                    hipEvent_t, hipEvent_t, const RangeArgs* ranges) {
    if (ranges && (ranges->tile_shift != 0 || !ranges->tile_range || begin_bit != 0)) return hipErrorInvalidValue;
    if (count > 1 && end_bit > begin_bit) {
        if (temp_bytes < radix_sort_temp_bytes(count, begin_bit, end_bit)) return hipErrorInvalidValue;
        const int passes = (end_bit - begin_bit + 7) / 8;
        uint64_t* k0 = keys[*sel];
        uint64_t* k1 = keys[*sel ^ 1];
        run_or_record(s, [=] {
            const uint32_t n = count_dev ? std::min<uint32_t>(*count_dev, (uint32_t)count) : (uint32_t)count;
            rd(k0, (size_t)n * 8);
            touch(k1, (size_t)n * 8);
        });
        if (passes & 1) *sel ^= 1;
    }
    if (!ranges) return hipSuccess;
    // the last pass (or, with no pass, the ranges kernel) ranges the sorted list
    RangeArgs r = *ranges;
    r.keys = keys[*sel];
    r.size = count;
    r.size_dev = count_dev;
    return launch_ranges(r, s);
}


// moves the data like an even (end-begin <= 32 bits: 4 passes) or odd number of passes would
// the words radix_sort_prepare zeroes (the histograms the mock sorts use), then the rest of the
// scratch as the look-back region: every span lies inside the scratch
hipError_t radix_sort_zero_spans(int32_t count, void* temp, size_t temp_bytes, FillSpan spans[2], uint32_t** hist) {
    if (count < 1) count = 1;
    if (temp_bytes < (size_t)count * 4 + 4096 || !temp || !hist) return hipErrorInvalidValue;
    *hist = static_cast<uint32_t*>(temp);
    const int64_t words = (int64_t)(temp_bytes / 4);
    const int64_t head = std::min<int64_t>(words, 8 * 256);
    spans[0] = {static_cast<uint32_t*>(temp), head, 0u};
    spans[1] = {static_cast<uint32_t*>(temp) + head, words - head, 0u};
    return hipSuccess;
}

hipError_t radix_sort_pairs_u32(uint32_t* keys[2], int32_t* values[2], int32_t* sel, int32_t count,
                                const uint32_t* count_dev, int32_t begin_bit, int32_t end_bit, void* temp,
                                size_t temp_bytes, hipStream_t s, bool) {
    return mock_sort(keys, values, sel, count, count_dev, begin_bit, end_bit, temp, temp_bytes, s);
}

template <typename KeyT>
hipError_t mock_sort(KeyT* keys[2], int32_t* values[2], int32_t* sel, int32_t count, const uint32_t* count_dev,
                     int32_t begin_bit, int32_t end_bit, void* temp, size_t temp_bytes, hipStream_t s) {
    if (count <= 1 || end_bit <= begin_bit) return hipSuccess;
    if (temp_bytes < radix_sort_temp_bytes(count, begin_bit, end_bit)) return hipErrorInvalidValue;
    const int passes = (end_bit - begin_bit + 7) / 8;
    KeyT* k0 = keys[*sel];
    KeyT* k1 = keys[*sel ^ 1];
    int32_t* v0 = values[*sel];
    int32_t* v1 = values[*sel ^ 1];
    const size_t tb = radix_sort_temp_bytes(count, begin_bit, end_bit);
    run_or_record(s, [=] {
        const uint32_t n = count_dev ? std::min<uint32_t>(*count_dev, (uint32_t)count) : (uint32_t)count;
        std::memcpy(k1, k0, (size_t)n * sizeof(KeyT));
        if (v0) std::memcpy(v1, v0, (size_t)n * 4);
        if (passes % 2 == 0) {
            std::memcpy(k0, k1, (size_t)n * sizeof(KeyT));
            if (v0) std::memcpy(v0, v1, (size_t)n * 4);
        }
        touch(temp, tb);
    });
    if (passes % 2) *sel ^= 1;
    return hipSuccess;
}

}  // namespace gr
