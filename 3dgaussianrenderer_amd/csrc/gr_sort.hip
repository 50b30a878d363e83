// gr_sort.hip -- stable LSD radix sort of (uint64 or uint32 key, int32 value)
// pairs for gfx950, replacing cub::DeviceRadixSort::SortPairs
// (core/gaussian_render.cu:591-632).
//
// Onesweep structure (Adinets & Merrill 2022, "Onesweep: A Faster Least
// Significant Digit Radix Sort for GPUs"), written for wave64:
//   init      zero the digit histograms, the partition tickets and the first
//             look-back region
//   histogram one read of all keys; 256-bin histograms of every 8-bit digit
//             (LDS-privatised, then global atomics)
//   scan      exclusive scan of each digit histogram -> global digit offsets
//   pass x P  each workgroup takes a partition ticket (in-order progress for
//             the look-back), loads kTile keys, ranks them per wave with
//             64-bit ballots (peer mask of equal digits, popcount of lower
//             lanes), publishes its per-digit counts, computes its global
//             per-digit prefix by decoupled look-back (Merrill & Garland 2016)
//             over the predecessors' status words, reorders the tile in LDS by
//             digit and writes it out digit-run by digit-run (coalesced).
// Stability: within a partition keys are ranked in their input order (wave
// by wave, item row by item row, lane by lane); partitions are ordered by
// ticket, so equal digits keep input order.  Status words are 32-bit:
// bits 30..31 flag (0 empty, 1 aggregate, 2 inclusive prefix), bits 0..29
// count, so one sort handles fewer than 2^30 pairs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "gr_kernels.hpp"

namespace gr {

// Partition shapes (lanes x pairs per lane) per sort: the 64-bit sorts (tile
// lists) and the 32-bit depth sort of the visible Gaussians; measured on the
// 6M-Gaussian 1080p frame (512 x 13 and 1024 x 8 fastest respectively; smaller
// partitions were slower for both).  The look-back reads 4 predecessors per
// step (16 or 32 per step measured slower).  Build-time overrides exist for
// sweeps.
#ifndef GR_SORT_THREADS
#define GR_SORT_THREADS 512
#endif
#ifndef GR_SORT_ITEMS
#define GR_SORT_ITEMS 13
#endif
#ifndef GR_DSORT_THREADS
#define GR_DSORT_THREADS 1024
#endif
#ifndef GR_DSORT_ITEMS
#define GR_DSORT_ITEMS 8
#endif
#ifndef GR_SORT_LOOKBACK
#define GR_SORT_LOOKBACK 4
#endif
// waves per SIMD the compiler is asked to fit (register budget); 0 = its own choice
#ifndef GR_SORT_WPE
#define GR_SORT_WPE 0
#endif
#ifndef GR_DSORT_WPE
#define GR_DSORT_WPE 0
#endif
template <int THREADS, int ITEMS, int WPE = 0>
struct SortShape {
    static constexpr int kThreads = THREADS;
    static constexpr int kWaves = THREADS / 64;
    static constexpr int kItems = ITEMS;
    static constexpr int kTile = THREADS * ITEMS;   // pairs per partition
    static constexpr int kWavesPerEU = WPE;
    static_assert(THREADS % 64 == 0 && THREADS >= 256, "one lane per digit in the look-back");
};
template <typename KeyT> struct ShapeOf { using type = SortShape<GR_SORT_THREADS, GR_SORT_ITEMS, GR_SORT_WPE>; };
template <> struct ShapeOf<uint32_t> { using type = SortShape<GR_DSORT_THREADS, GR_DSORT_ITEMS, GR_DSORT_WPE>; };
constexpr int kMinTile = GR_SORT_THREADS * GR_SORT_ITEMS < GR_DSORT_THREADS * GR_DSORT_ITEMS
                             ? GR_SORT_THREADS * GR_SORT_ITEMS
                             : GR_DSORT_THREADS * GR_DSORT_ITEMS;
constexpr int kRadix = 256;
constexpr int kMaxPasses = 8;
// copies of the digit histograms the histogram read's workgroups (and the emit's) flush into
// (workgroup b into copy b mod kHistReplicas, fewer atomics per address); the scan sums them
// into copy 0
constexpr int kHistReplicas = kSortHistCopies;
static_assert(kSortHistCopyStride == kMaxPasses * kRadix, "histogram copy layout");
constexpr int kLookback = GR_SORT_LOOKBACK;            // predecessors read per look-back step
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagIncl = 2u << 30;
constexpr uint32_t kCountMask = (1u << 30) - 1u;

// Scratch layout: histograms, tickets, two look-back regions sized for the
// smallest partition shape (so one scratch serves every sort of `count` pairs);
// `partitions` is the count of the sort at hand.
struct SortLayout {
    size_t hist_off, ticket_off, status_off[2], total;
    int partitions;
};

static SortLayout sort_layout(int32_t count, int tile) {
    SortLayout l;
    l.partitions = (count + tile - 1) / tile;
    const int64_t max_parts = ((int64_t)count + kMinTile - 1) / kMinTile;
    size_t off = 0;
    l.hist_off = off;
    off += sizeof(uint32_t) * kHistReplicas * kMaxPasses * kRadix;
    l.ticket_off = off;   // kMaxPasses tickets, then kMaxPasses trivial-pass flags at +128
    off += 256;
    off = (off + 255) & ~(size_t)255;
    const size_t region = (size_t)max_parts * kRadix * sizeof(uint32_t);
    l.status_off[0] = off;
    off += (region + 255) & ~(size_t)255;
    l.status_off[1] = off;
    off += (region + 255) & ~(size_t)255;
    l.total = off;
    return l;
}

template <typename KeyT> static SortLayout layout_for(int32_t count) {
    return sort_layout(count < 1 ? 1 : count, ShapeOf<KeyT>::type::kTile);
}

size_t radix_sort_temp_bytes(int32_t count, int32_t, int32_t) {
    return sort_layout(count < 1 ? 1 : count, kMinTile).total;
}

__global__ void sort_init_kernel(uint32_t* hist, uint32_t* tickets, uint32_t* status0, int64_t status_n) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < kHistReplicas * kMaxPasses * kRadix) hist[tid] = 0u;
    if (tid < kMaxPasses) tickets[tid] = 0u;
    for (int64_t i = tid; i < status_n; i += (int64_t)gridDim.x * blockDim.x) status0[i] = 0u;
}

// Digit histograms of every pass in one read.  A workgroup stages 4096 keys
// in LDS with coalesced loads; each lane then walks 16 consecutive keys and
// adds runs of equal digits with one LDS atomic per run.  Tile-list keys
// arrive in emission order, where neighbours share digits, so a lane per key
// would put many lanes of one instruction on one LDS address.  The staging
// array is skewed by one slot per 16 keys to spread the per-lane walks over
// the LDS banks.
#ifndef GR_HIST_BLOCKS
#define GR_HIST_BLOCKS 1024
#endif

#ifndef GR_HIST_SEG
#define GR_HIST_SEG 16
#endif

constexpr int kHistBlocks = GR_HIST_BLOCKS;   // workgroups of the histogram read (each flushes once)
constexpr int kHistThreads = 256;
constexpr int kHistSeg = GR_HIST_SEG;         // consecutive keys per lane
constexpr int kHistChunk = kHistThreads * kHistSeg;

template <typename KeyT>
__global__ __launch_bounds__(kHistThreads) void sort_hist_kernel(const KeyT* __restrict__ keys, int32_t n,
                                                                 const uint32_t* count_dev, int begin_bit,
                                                                 int end_bit, int passes, uint32_t* hist) {
    constexpr int kP = (int)sizeof(KeyT);   // at most one 8-bit pass per key byte
    __shared__ uint32_t s_h[kP][kRadix];
    const int dw = (end_bit - begin_bit + passes - 1) / passes;   // the passes' digit width
    if (count_dev) n = (int32_t)min((uint32_t)n, *count_dev);
    __shared__ KeyT s_k[kHistChunk + kHistChunk / kHistSeg];
    for (int i = threadIdx.x; i < kP * kRadix; i += blockDim.x) (&s_h[0][0])[i] = 0u;
    for (int64_t base = (int64_t)blockIdx.x * kHistChunk; base < n; base += (int64_t)gridDim.x * kHistChunk) {
        const int m = (int)min((int64_t)kHistChunk, (int64_t)n - base);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kHistSeg; ++k) {
            const int i = k * kHistThreads + threadIdx.x;
            if (i < m) s_k[i + i / kHistSeg] = keys[base + i];
        }
        __syncthreads();
        uint32_t cur[kP], cnt[kP];
#pragma unroll
        for (int p = 0; p < kP; ++p) {
            cur[p] = 0u;
            cnt[p] = 0u;
        }
        const int my = threadIdx.x * kHistSeg;
        for (int j = 0; j < kHistSeg && my + j < m; ++j) {
            const KeyT k = s_k[my + j + threadIdx.x];
#pragma unroll
            for (int p = 0; p < kP; ++p) {
                if (p < passes) {
                    const int shift = begin_bit + dw * p;
                    const int bits = min(dw, end_bit - shift);
                    const uint32_t d = (uint32_t)(k >> shift) & ((1u << bits) - 1u);
                    if (d != cur[p]) {
                        if (cnt[p]) atomicAdd(&s_h[p][cur[p]], cnt[p]);
                        cur[p] = d;
                        cnt[p] = 0u;
                    }
                    ++cnt[p];
                }
            }
        }
#pragma unroll
        for (int p = 0; p < kP; ++p)
            if (p < passes && cnt[p]) atomicAdd(&s_h[p][cur[p]], cnt[p]);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * kRadix; i += blockDim.x) {
        const uint32_t v = (&s_h[0][0])[i];
        if (v) atomicAdd(&hist[(blockIdx.x % kHistReplicas) * kMaxPasses * kRadix + i], v);
    }
}

// exclusive scan of each pass's 256 bins; one workgroup of 256 lanes per pass.  A pass whose
// keys all share one digit is flagged trivial (trivial[pass] = 1): its stable permutation is the
// identity, so the pass copies its partitions through unchanged.
__global__ __launch_bounds__(256) void sort_hist_scan_kernel(uint32_t* hist, uint32_t* trivial) {
    __shared__ uint32_t s_w[4];
    uint32_t* h = hist + blockIdx.x * kRadix;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    uint32_t v = h[t];
    for (int r = 1; r < kHistReplicas; ++r) v += h[(size_t)r * kMaxPasses * kRadix + t];
    uint32_t incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(incl, d, 64);
        if (lane >= d) incl += o;
    }
    if (lane == 63) s_w[w] = incl;
    __syncthreads();
    uint32_t off = 0;
    for (int k = 0; k < w; ++k) off += s_w[k];
    h[t] = off + incl - v;
    const uint32_t total = s_w[0] + s_w[1] + s_w[2] + s_w[3];
    const int one_digit = __syncthreads_or(total > 0u && v == total);
    if (t == 0) trivial[blockIdx.x] = one_digit ? 1u : 0u;
}

__device__ __forceinline__ uint32_t ld_status(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Publishes partition `part`'s count of digit t: an aggregate, or for partition 0 already the
// inclusive prefix.
__device__ __forceinline__ void lb_publish(uint32_t* status, int part, int t, uint32_t total) {
    st_status(status + (size_t)part * kRadix + t, (part == 0 ? kFlagIncl : kFlagAgg) | total);
}

// Returns the count of digit t in all partitions before `part` (decoupled look-back over
// kLookback predecessors per step: aggregates are summed until an inclusive prefix is found) and
// publishes part's inclusive prefix.  Every wait is on an earlier partition.
__device__ __forceinline__ uint32_t lb_resolve(uint32_t* status, int part, int t, uint32_t total) {
    if (part == 0) return 0u;
    uint32_t excl = 0u;
    int j = part - 1;
    bool done = false;
    while (!done) {
        uint32_t sv[kLookback];
#pragma unroll
        for (int q = 0; q < kLookback; ++q)
            sv[q] = (j - q >= 0) ? ld_status(status + (size_t)(j - q) * kRadix + t) : kFlagIncl;
        int q = 0;
        for (; q < kLookback; ++q) {
            const uint32_t flag = sv[q] & ~kCountMask;
            if (flag == 0u) break;               // not yet published: poll again from here
            excl += sv[q] & kCountMask;
            if (flag == kFlagIncl) {
                done = true;
                break;
            }
        }
        if (!done) j -= q;
    }
    st_status(status + (size_t)part * kRadix + t, kFlagIncl | (excl + total));
    return excl;
}

// Stable rank of a wave's keys among the wave's earlier keys of the same digit, item row by item
// row: every lane ORs its bit into its digit's 64-bit lane mask in LDS (OR commutes, so the mask
// does not depend on the order the LDS applies the lanes), reads the mask back -- its peers --
// and the row's first lane of a digit adds the row's count to the digit's counter and clears the
// mask.  rank[i] = wave-local rank; counters and masks start at zero and the counters end as the
// wave's digit counts.  (Against 64-bit ballots per digit bit: fewer VALU operations per key and
// a shorter ranking, which shortens the look-back chain.)
template <typename KeyT, int IT, int RW>
__device__ __forceinline__ void wave_rank(const KeyT (&key)[IT], uint32_t (&rank)[IT], int64_t wb, int32_t n,
                                          int shift, uint32_t dmask, uint32_t* wcnt, uint64_t* mask, int lane) {
    const uint64_t lt = (1ull << lane) - 1ull, me = 1ull << lane;
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const bool valid = wb + i * 64 < n;
        const uint32_t d = (uint32_t)(key[i] >> shift) & dmask;
        if (valid) __hip_atomic_fetch_or(&mask[d], me, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const uint64_t peers =
            valid ? __hip_atomic_load(&mask[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) : 0ull;
        const uint32_t below = (uint32_t)__popcll(peers & lt);
        const uint32_t prev = wcnt[d];
        rank[i] = prev + below;
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (valid && below == 0u) {
            wcnt[d] = prev + (uint32_t)__popcll(peers);
            __hip_atomic_store(&mask[d], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

// Digit counts of a ranked partition (threads < RW own one digit each; wcnt[k][d] becomes the
// exclusive prefix over the waves), an exclusive scan of them over the digits into s_dstart (all
// kRadix entries; digits without keys are 0), and the count published for the look-back.  Needs
// a barrier before (the ranking) and has two inside; returns the digit's total and start.
template <int WV, int RW>
__device__ __forceinline__ void partition_digits(uint32_t (*wcnt)[RW], uint32_t* s_dstart, uint32_t* s_wsum,
                                                 uint32_t* status, uint32_t* status_next, int part, int nd, int t,
                                                 uint32_t& total, uint32_t& dstart) {
    const int lane = t & 63, w = t >> 6;
    total = 0u;
    if (t < kRadix) {
        if (t < RW) {
#pragma unroll
            for (int k = 0; k < WV; ++k) {
                const uint32_t c = wcnt[k][t];
                wcnt[k][t] = total;
                total += c;
            }
        }
        if (t < nd) lb_publish(status, part, t, total);
        if (status_next) status_next[(size_t)part * kRadix + t] = 0u;
        uint32_t incl = total;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(incl, d, 64);
            if (lane >= d) incl += o;
        }
        if (lane == 63) s_wsum[w] = incl;
        s_dstart[t] = incl - total;    // within-wave part, completed below
    }
    __syncthreads();
    if (t < kRadix) {
        uint32_t woff = 0;
        for (int k = 0; k < w; ++k) woff += s_wsum[k];
        dstart = s_dstart[t] + woff;
        s_dstart[t] = dstart;
    }
    __syncthreads();
}

// Pairs passes (the depth sort, the stage-level pair sort): one partition per workgroup.  The
// partition's counts are published right after the ranking, and its look-back is resolved after
// the reorder in LDS, so the wait overlaps the local work.  The per-wave digit counters and lane
// masks are dead once every pair's position in the partition is known, so the reordered
// partition overlays them.
template <typename KeyT, bool PAIRS, class S>
struct SortLds {
    static constexpr size_t kPairBytes = (size_t)S::kTile * (sizeof(KeyT) + (PAIRS ? sizeof(int32_t) : 0));
    static constexpr size_t kCountBytes = (size_t)S::kWaves * kRadix * (sizeof(uint32_t) + sizeof(uint64_t));
    static constexpr size_t kBytes = kPairBytes > kCountBytes ? kPairBytes : kCountBytes;
};

template <typename KeyT, bool PAIRS, class S>
__global__ __launch_bounds__(S::kThreads) __attribute__((amdgpu_waves_per_eu(S::kWavesPerEU))) void onesweep_kernel(
    const KeyT* __restrict__ kin, const int32_t* __restrict__ vin, KeyT* __restrict__ kout,
    int32_t* __restrict__ vout, int32_t n, const uint32_t* count_dev, int strict, int shift, int bits,
    const uint32_t* __restrict__ digit_off, uint32_t* status, uint32_t* status_next, uint32_t* ticket,
    const uint32_t* trivial) {
    constexpr int kSortThreads = S::kThreads;
    constexpr int kSortWaves = S::kWaves;
    constexpr int kSortItems = S::kItems;
    constexpr int kSortTile = S::kTile;
    __shared__ uint32_t s_part;
    __shared__ uint32_t s_dstart[kRadix];
    __shared__ uint32_t s_gbase[kRadix];
    __shared__ uint32_t s_wsum[kRadix / 64];
    __shared__ __align__(16) unsigned char s_raw[SortLds<KeyT, PAIRS, S>::kBytes];
    auto s_wcnt = reinterpret_cast<uint32_t(*)[kRadix]>(s_raw);                        // ranking phase
    auto s_mask = reinterpret_cast<uint64_t(*)[kRadix]>(s_raw + (size_t)kSortWaves * kRadix * 4);
    KeyT* s_keys = reinterpret_cast<KeyT*>(s_raw);                                       // reorder phase
    int32_t* s_vals = reinterpret_cast<int32_t*>(s_raw + (size_t)kSortTile * sizeof(KeyT));

    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) s_part = atomicAdd(ticket, 1u);
    for (int i = t; i < kSortWaves * kRadix; i += kSortThreads) {
        (&s_wcnt[0][0])[i] = 0u;
        (&s_mask[0][0])[i] = 0ull;
    }
    __syncthreads();
    const int part = (int)s_part;
    if (count_dev) {
        // strict (histograms made by the producer): a count above the capacity means the list
        // was not written -- sort nothing
        if (strict && *count_dev > (uint32_t)n) return;
        n = (int32_t)min((uint32_t)n, *count_dev);
    }
    if ((int64_t)part * kSortTile >= n) return;   // beyond the device-side count: whole workgroup

    const int64_t base = (int64_t)part * kSortTile;
    if (*trivial) {
        // every key has the same digit: the pass is the identity, the partition is copied through
        if (status_next && t < kRadix) status_next[(size_t)part * kRadix + t] = 0u;
        const int64_t end = min((int64_t)n, base + kSortTile);
        for (int64_t i = base + t; i < end; i += kSortThreads) {
            kout[i] = kin[i];
            if constexpr (PAIRS) vout[i] = vin[i];
        }
        return;
    }
    const int64_t wbase = base + (int64_t)w * (64 * kSortItems) + lane;
    const uint32_t dmask = (1u << bits) - 1u;
    const int nd = 1 << bits;
    KeyT key[kSortItems];
    int32_t val[kSortItems];
    uint32_t rank[kSortItems];
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        const int64_t idx = wbase + i * 64;
        key[i] = idx < n ? kin[idx] : (KeyT)0;
        if constexpr (PAIRS) val[i] = idx < n ? vin[idx] : 0;
    }
    wave_rank<KeyT, kSortItems, kRadix>(key, rank, wbase, n, shift, dmask, s_wcnt[w], s_mask[w], lane);
    __syncthreads();
    uint32_t total = 0u, dstart = 0u;
    partition_digits<kSortWaves, kRadix>(s_wcnt, s_dstart, s_wsum, status, status_next, part, nd, t, total, dstart);

    // positions in the partition, then (counters dead) the reorder over the same LDS
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        const uint32_t d = (uint32_t)(key[i] >> shift) & dmask;
        rank[i] += s_dstart[d] + s_wcnt[w][d];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        if (wbase + i * 64 < n) {
            s_keys[rank[i]] = key[i];
            if constexpr (PAIRS) s_vals[rank[i]] = val[i];
        }
    }
    if (t < nd) s_gbase[t] = digit_off[t] + lb_resolve(status, part, t, total) - dstart;
    __syncthreads();
    const int nvalid = (int)min((int64_t)kSortTile, (int64_t)n - base);
    if (nvalid == kSortTile) {
        // a whole partition: the lane's LDS reads are issued together (one round trip per step
        // instead of two per entry)
        KeyT kk[kSortItems];
        uint32_t gp[kSortItems];
#pragma unroll
        for (int i = 0; i < kSortItems; ++i) kk[i] = s_keys[t + i * kSortThreads];
#pragma unroll
        for (int i = 0; i < kSortItems; ++i)
            gp[i] = s_gbase[(uint32_t)(kk[i] >> shift) & dmask] + (uint32_t)(t + i * kSortThreads);
#pragma unroll
        for (int i = 0; i < kSortItems; ++i) {
            kout[gp[i]] = kk[i];
            if constexpr (PAIRS) vout[gp[i]] = s_vals[t + i * kSortThreads];
        }
        return;
    }
    for (int j = t; j < nvalid; j += kSortThreads) {
        const KeyT k = s_keys[j];
        const uint32_t d = (uint32_t)(k >> shift) & dmask;
        const uint32_t gpos = s_gbase[d] + (uint32_t)j;
        kout[gpos] = k;
        if constexpr (PAIRS) vout[gpos] = s_vals[j];
    }
}

// Keys-only passes (tile lists): persistent workgroups that hold two partitions each.  A
// workgroup takes partition Y while X -- ranked, counts published, reordered in LDS -- still
// waits for its look-back; the waves that own X's digits resolve X's look-back while the other
// waves rank their rows of Y, then rank theirs; Y's counts are published, X is stored and Y is
// reordered into the LDS X held.  Progress: every look-back waits on earlier partitions only, and
// the owner of the earliest unresolved partition never waits on an unresolved one, so it
// finishes it.  The digit tables are RW = 2^RB wide (7-bit digits fit more keys per partition
// in the LDS that keeps two workgroups on a CU).
//
// Tile ranges (ranges != NULL, last pass of a tile-list sort: the key's low 32 bits are the tile
// id, which holds every sorted bit): the output is sorted by tile, and the entries of one
// partition's digit run sit contiguously in the output in LDS order, so an entry whose LDS
// neighbour has another tile (or none) is the first or last of its tile within that run.  Such
// an entry lowers its tile's start to its position (atomicMin, as unsigned: the ranges hold -1
// beforehand) or raises its tile's end to its position + 1 (atomicMax); the runs of one tile in
// several partitions are adjacent in the output, so the minimum and maximum are the tile's
// range.  Tiles without entries keep -1.  Every write is an atomic, so writes to one tile from
// workgroups on different XCDs combine.
template <int THREADS, int ITEMS, int RB>
struct KeysShape {
    static constexpr int kThreads = THREADS;
    static constexpr int kWaves = THREADS / 64;
    static constexpr int kItems = ITEMS;
    static constexpr int kTile = THREADS * ITEMS;
    static constexpr int kRadixBits = RB;
    static_assert(THREADS % 64 == 0 && THREADS >= 256, "one lane per digit in the look-back");
};
#ifndef GR_KSORT_ITEMS7
#define GR_KSORT_ITEMS7 15
#endif
#ifndef GR_KSORT_ITEMS8
#define GR_KSORT_ITEMS8 13
#endif
using KeysShape7 = KeysShape<512, GR_KSORT_ITEMS7, 7>;
using KeysShape8 = KeysShape<512, GR_KSORT_ITEMS8, 8>;
static_assert(KeysShape7::kTile >= kMinTile && KeysShape8::kTile >= kMinTile,
              "the look-back regions hold count / kMinTile partitions");

// Entries per lane whose LDS reads the store of a whole partition issues together.
#ifndef GR_KSORT_STORE_GROUP
#define GR_KSORT_STORE_GROUP 3
#endif
constexpr int kStoreGroup = GR_KSORT_STORE_GROUP;

template <class S>
__global__ __launch_bounds__(S::kThreads) __attribute__((amdgpu_waves_per_eu(4))) void onesweep_keys_kernel(
    const uint64_t* __restrict__ kin, uint64_t* __restrict__ kout, int32_t n, const uint32_t* count_dev, int strict,
    int shift, int bits, const uint32_t* __restrict__ digit_off, uint32_t* status, uint32_t* status_next,
    uint32_t* ticket, const uint32_t* trivial, int32_t* ranges, int32_t ntiles) {
    constexpr int T = S::kThreads, WV = S::kWaves, IT = S::kItems, TILE = S::kTile;
    constexpr int RW = 1 << S::kRadixBits;
    __shared__ uint32_t s_part;
    __shared__ uint32_t s_dstart[kRadix];
    __shared__ uint32_t s_gbase[kRadix];
    __shared__ uint32_t s_wsum[kRadix / 64];
    __shared__ uint32_t s_wcnt[WV][RW];
    __shared__ uint64_t s_mask[WV][RW];
    __shared__ uint64_t s_keys[TILE];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (count_dev) {
        if (strict && *count_dev > (uint32_t)n) return;
        n = (int32_t)min((uint32_t)n, *count_dev);
    }
    const int parts = (int)(((int64_t)n + TILE - 1) / TILE);
    const uint32_t dmask = (1u << bits) - 1u;
    const int nd = 1 << bits;
    auto take = [&]() -> int {
        if (t == 0) s_part = atomicAdd(ticket, 1u);
        __syncthreads();
        const int p = (int)s_part;
        __syncthreads();
        return p;
    };
    int X = take();
    if (X >= parts) return;
    if (*trivial) {
        // every key has the same digit: the pass is the identity, partitions are copied through
        // (ranges: each boundary is found against the neighbours in the list itself, one plain
        // store per start and end)
        for (;;) {
            const int64_t base = (int64_t)X * TILE, end = min((int64_t)n, base + TILE);
            if (status_next && t < kRadix) status_next[(size_t)X * kRadix + t] = 0u;
            for (int64_t i = base + t; i < end; i += T) {
                const uint64_t k = kin[i];
                kout[i] = k;
                const uint32_t tile = (uint32_t)k;
                if (ranges && tile < (uint32_t)ntiles) {
                    if (i == 0 || (uint32_t)kin[i - 1] != tile) ranges[2 * (size_t)tile] = (int32_t)i;
                    if (i + 1 == n || (uint32_t)kin[i + 1] != tile) ranges[2 * (size_t)tile + 1] = (int32_t)(i + 1);
                }
            }
            X = take();
            if (X >= parts) return;
        }
    }
    uint64_t key[IT];
    uint32_t rank[IT];
    auto wbase = [&](int P) { return (int64_t)P * TILE + (int64_t)w * (64 * IT) + lane; };
    auto load = [&](int P) {
        const int64_t wb = wbase(P);
#pragma unroll
        for (int i = 0; i < IT; ++i) key[i] = wb + i * 64 < n ? kin[wb + i * 64] : 0ull;
    };
    auto zero_tables = [&]() {
        for (int i = t; i < WV * RW; i += T) {
            (&s_wcnt[0][0])[i] = 0u;
            (&s_mask[0][0])[i] = 0ull;
        }
    };
    // after wave_rank and a barrier: publishes P's counts, leaves the partition-local positions
    // in rank[]
    auto finish = [&](int P, uint32_t& total, uint32_t& dstart) {
        partition_digits<WV, RW>(s_wcnt, s_dstart, s_wsum, status, status_next, P, nd, t, total, dstart);
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const uint32_t d = (uint32_t)(key[i] >> shift) & dmask;
            rank[i] += s_dstart[d] + s_wcnt[w][d];
        }
    };
    auto reorder = [&](int P) {
        const int64_t wb = wbase(P);
#pragma unroll
        for (int i = 0; i < IT; ++i)
            if (wb + i * 64 < n) s_keys[rank[i]] = key[i];
    };
    uint32_t totX = 0u, dsX = 0u;
    load(X);
    zero_tables();
    __syncthreads();
    wave_rank<uint64_t, IT, RW>(key, rank, wbase(X), n, shift, dmask, s_wcnt[w], s_mask[w], lane);
    __syncthreads();
    finish(X, totX, dsX);
    reorder(X);
    for (;;) {
        const int Y = take();   // its barriers also order the reorder of X before the store
        const bool haveY = Y < parts;
        uint32_t totY = 0u, dsY = 0u;
        if (haveY) {
            load(Y);
            zero_tables();
        }
        __syncthreads();
        if (t < nd) s_gbase[t] = digit_off[t] + lb_resolve(status, X, t, totX) - dsX;
        if (haveY) wave_rank<uint64_t, IT, RW>(key, rank, wbase(Y), n, shift, dmask, s_wcnt[w], s_mask[w], lane);
        __syncthreads();
        if (haveY) finish(Y, totY, dsY);
        const int nvalid = (int)min((int64_t)TILE, (int64_t)n - (int64_t)X * TILE);
        // entry j of the reordered partition goes to its digit's base + j; with ranges, equal tiles
        // have equal digits, so a neighbour with the same tile is in the same run (neighbours'
        // tiles read clamped to the partition)
        const uint32_t* s_lo = reinterpret_cast<const uint32_t*>(s_keys);   // low words: the tiles
        auto store = [&](int j, uint64_t k, uint32_t o, uint32_t tp, uint32_t tn) {
            kout[o] = k;
            const uint32_t tile = (uint32_t)k;
            if (ranges && tile < (uint32_t)ntiles) {
                if (j == 0 || tp != tile) atomicMin(reinterpret_cast<uint32_t*>(ranges) + 2 * (size_t)tile, o);
                if (j + 1 == nvalid || tn != tile) atomicMax(ranges + 2 * (size_t)tile + 1, (int32_t)(o + 1u));
            }
        };
        if (nvalid == TILE) {
            // a whole partition: the LDS reads of kStoreGroup entries per lane are issued together
            // (one LDS round trip per group instead of two per entry)
#pragma unroll
            for (int i0 = 0; i0 < IT; i0 += kStoreGroup) {
                uint64_t kk[kStoreGroup];
                uint32_t oo[kStoreGroup], tp[kStoreGroup], tn[kStoreGroup];
#pragma unroll
                for (int u = 0; u < kStoreGroup; ++u) {
                    if (i0 + u < IT) {
                        const int j = t + (i0 + u) * T;
                        kk[u] = s_keys[j];
                        if (ranges) {
                            tp[u] = s_lo[2 * (j > 0 ? j - 1 : j)];
                            tn[u] = s_lo[2 * (j + 1 < TILE ? j + 1 : j)];
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < kStoreGroup; ++u)
                    if (i0 + u < IT) oo[u] = s_gbase[(uint32_t)(kk[u] >> shift) & dmask] + (uint32_t)(t + (i0 + u) * T);
#pragma unroll
                for (int u = 0; u < kStoreGroup; ++u)
                    if (i0 + u < IT) store(t + (i0 + u) * T, kk[u], oo[u], ranges ? tp[u] : 0u, ranges ? tn[u] : 0u);
            }
        } else {
            for (int j = t; j < nvalid; j += T) {
                const uint64_t k = s_keys[j];
                const uint32_t o = s_gbase[(uint32_t)(k >> shift) & dmask] + (uint32_t)j;
                store(j, k, o, ranges ? s_lo[2 * (j > 0 ? j - 1 : j)] : 0u,
                      ranges ? s_lo[2 * (j + 1 < nvalid ? j + 1 : j)] : 0u);
            }
        }
        if (!haveY) return;
        __syncthreads();
        reorder(Y);
        X = Y;
        totX = totY;
        dsX = dsY;
    }
}

// Resident workgroups of the keys kernel on the current device (the persistent grid).
template <class S>
static int keys_grid() {
    static std::mutex mu;
    static int grid[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    std::lock_guard<std::mutex> lk(mu);
    if (!grid[dev]) {
        int cus = 0, per = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)onesweep_keys_kernel<S>, S::kThreads, 0) !=
                hipSuccess || per <= 0)
            per = 1;
        grid[dev] = cus * per;
    }
    return grid[dev];
}

static int sort_passes(int32_t begin_bit, int32_t end_bit) { return (end_bit - begin_bit + 7) / 8; }

int sort_digit_bits(int32_t begin_bit, int32_t end_bit) {
    const int p = sort_passes(begin_bit, end_bit);
    return p > 0 ? (end_bit - begin_bit + p - 1) / p : 0;
}

// Zeroes the histograms, the partition tickets and the first look-back region.
static hipError_t sort_prepare(const SortLayout& l, char* tb, hipStream_t s) {
    uint32_t* hist = reinterpret_cast<uint32_t*>(tb + l.hist_off);
    uint32_t* tickets = reinterpret_cast<uint32_t*>(tb + l.ticket_off);
    uint32_t* status0 = reinterpret_cast<uint32_t*>(tb + l.status_off[0]);
    const int64_t status_n = (int64_t)l.partitions * kRadix;
    int64_t init_blocks = (status_n + 255) / 256;
    if (init_blocks < (kHistReplicas * kMaxPasses * kRadix + 255) / 256)
        init_blocks = (kHistReplicas * kMaxPasses * kRadix + 255) / 256;
    if (init_blocks > 4096) init_blocks = 4096;
    hipLaunchKernelGGL(sort_init_kernel, dim3((unsigned)init_blocks), dim3(256), 0, s, hist, tickets, status0,
                       status_n);
    return hipGetLastError();
}

template <typename KeyT>
static hipError_t sort_check(int32_t count, int32_t begin_bit, int32_t end_bit, const void* temp, size_t temp_bytes) {
    constexpr int kKeyBits = 8 * (int)sizeof(KeyT);
    if ((uint32_t)count > kCountMask || begin_bit < 0 || end_bit > kKeyBits) return hipErrorInvalidValue;
    if (sort_passes(begin_bit, end_bit) > kMaxPasses) return hipErrorInvalidValue;
    if (temp_bytes < layout_for<KeyT>(count).total || temp == nullptr) return hipErrorInvalidValue;
    return hipSuccess;
}

// Histogram scan + the onesweep passes (histograms already accumulated).
template <typename KeyT>
static hipError_t sort_passes_t(KeyT* keys[2], int32_t* values[2], int32_t* sel, int32_t count,
                                const uint32_t* count_dev, int32_t begin_bit, int32_t end_bit, char* tb, int strict,
                                hipStream_t s, hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr,
                                int32_t* ranges = nullptr, int32_t ntiles = 0) {
    using S = typename ShapeOf<KeyT>::type;
    const SortLayout l = layout_for<KeyT>(count);
    const int passes = sort_passes(begin_bit, end_bit);
    uint32_t* hist = reinterpret_cast<uint32_t*>(tb + l.hist_off);
    uint32_t* tickets = reinterpret_cast<uint32_t*>(tb + l.ticket_off);
    uint32_t* trivial = reinterpret_cast<uint32_t*>(tb + l.ticket_off + 128);
    uint32_t* status[2] = {reinterpret_cast<uint32_t*>(tb + l.status_off[0]),
                           reinterpret_cast<uint32_t*>(tb + l.status_off[1])};
    hipLaunchKernelGGL(sort_hist_scan_kernel, dim3(passes), dim3(256), 0, s, hist, trivial);
    if (ev_begin) {
        const hipError_t e = hipEventRecord(ev_begin, s);
        if (e != hipSuccess) return e;
    }
    const int dw = sort_digit_bits(begin_bit, end_bit);
    for (int p = 0; p < passes; ++p) {
        const int shift = begin_bit + dw * p;
        const int bits = min(dw, end_bit - shift);
        const int cur = *sel;
        uint32_t* st = status[p & 1];
        uint32_t* st_next = (p + 1 < passes) ? status[(p + 1) & 1] : nullptr;
        if (values[0]) {
            hipLaunchKernelGGL((onesweep_kernel<KeyT, true, S>), dim3(l.partitions), dim3(S::kThreads), 0, s,
                               keys[cur], values[cur], keys[cur ^ 1], values[cur ^ 1], count, count_dev, strict,
                               shift, bits, hist + p * kRadix, st, st_next, tickets + p, trivial + p);
        } else if constexpr (sizeof(KeyT) == 8) {
            // keys only: the persistent kernel, with 7-bit digit tables where the digit fits them;
            // the last pass also writes the tile ranges when asked
            int32_t* rg = (p + 1 == passes) ? ranges : nullptr;
            if (bits <= 7) {
                const int parts = (count + KeysShape7::kTile - 1) / KeysShape7::kTile;
                hipLaunchKernelGGL(onesweep_keys_kernel<KeysShape7>, dim3(std::min(keys_grid<KeysShape7>(), parts)),
                                   dim3(KeysShape7::kThreads), 0, s, keys[cur], keys[cur ^ 1], count, count_dev,
                                   strict, shift, bits, hist + p * kRadix, st, st_next, tickets + p, trivial + p,
                                   rg, ntiles);
            } else {
                const int parts = (count + KeysShape8::kTile - 1) / KeysShape8::kTile;
                hipLaunchKernelGGL(onesweep_keys_kernel<KeysShape8>, dim3(std::min(keys_grid<KeysShape8>(), parts)),
                                   dim3(KeysShape8::kThreads), 0, s, keys[cur], keys[cur ^ 1], count, count_dev,
                                   strict, shift, bits, hist + p * kRadix, st, st_next, tickets + p, trivial + p,
                                   rg, ntiles);
            }
        } else {
            return hipErrorInvalidValue;   // keys-only sorts are 64-bit
        }
        *sel = cur ^ 1;
    }
    if (ev_end) {
        const hipError_t e = hipEventRecord(ev_end, s);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

template <typename KeyT>
static hipError_t sort_pairs_t(KeyT* keys[2], int32_t* values[2], int32_t* sel, int32_t count,
                               const uint32_t* count_dev, int32_t begin_bit, int32_t end_bit, void* temp,
                               size_t temp_bytes, hipStream_t s, bool zeroed = false) {
    if (count <= 1 || end_bit <= begin_bit) return hipSuccess;
    hipError_t e = sort_check<KeyT>(count, begin_bit, end_bit, temp, temp_bytes);
    if (e != hipSuccess) return e;
    const SortLayout l = layout_for<KeyT>(count);
    char* tb = static_cast<char*>(temp);
    if (!zeroed) {
        e = sort_prepare(l, tb, s);
        if (e != hipSuccess) return e;
    }
    int hist_blocks = (count + kHistChunk - 1) / kHistChunk;
    if (hist_blocks > kHistBlocks) hist_blocks = kHistBlocks;
    hipLaunchKernelGGL(sort_hist_kernel<KeyT>, dim3(hist_blocks), dim3(kHistThreads), 0, s, keys[*sel], count,
                       count_dev, begin_bit, end_bit, sort_passes(begin_bit, end_bit),
                       reinterpret_cast<uint32_t*>(tb + l.hist_off));
    return sort_passes_t<KeyT>(keys, values, sel, count, count_dev, begin_bit, end_bit, tb, 0, s);
}

hipError_t radix_sort_pairs(uint64_t* keys[2], int32_t* values[2], int32_t* sel, int32_t count,
                            const uint32_t* count_dev, int32_t begin_bit, int32_t end_bit, void* temp,
                            size_t temp_bytes, hipStream_t s) {
    return sort_pairs_t<uint64_t>(keys, values, sel, count, count_dev, begin_bit, end_bit, temp, temp_bytes, s);
}

hipError_t radix_sort_pairs_u32(uint32_t* keys[2], int32_t* values[2], int32_t* sel, int32_t count,
                                const uint32_t* count_dev, int32_t begin_bit, int32_t end_bit, void* temp,
                                size_t temp_bytes, hipStream_t s, bool zeroed) {
    return sort_pairs_t<uint32_t>(keys, values, sel, count, count_dev, begin_bit, end_bit, temp, temp_bytes, s,
                                  zeroed);
}

// The words sort_prepare zeroes: the histogram copies and the partition tickets (one span), and
// the first look-back region's slots of the sort's partitions.
static void zero_spans(const SortLayout& l, char* tb, FillSpan spans[2]) {
    spans[0] = {reinterpret_cast<uint32_t*>(tb + l.hist_off),
                (int64_t)((l.ticket_off + kMaxPasses * sizeof(uint32_t) - l.hist_off) / sizeof(uint32_t)), 0u};
    spans[1] = {reinterpret_cast<uint32_t*>(tb + l.status_off[0]), (int64_t)l.partitions * kRadix, 0u};
}

hipError_t radix_sort_zero_spans(int32_t count, void* temp, size_t temp_bytes, FillSpan spans[2], uint32_t** hist) {
    const SortLayout l = layout_for<uint64_t>(count);
    if (temp_bytes < l.total || temp == nullptr || hist == nullptr || spans == nullptr) return hipErrorInvalidValue;
    *hist = reinterpret_cast<uint32_t*>(static_cast<char*>(temp) + l.hist_off);
    zero_spans(l, static_cast<char*>(temp), spans);
    return hipSuccess;
}

hipError_t radix_sort_prepare(int32_t count, void* temp, size_t temp_bytes, uint32_t** hist, hipStream_t s) {
    const SortLayout l = layout_for<uint64_t>(count);
    if (temp_bytes < l.total || temp == nullptr || hist == nullptr) return hipErrorInvalidValue;
    *hist = reinterpret_cast<uint32_t*>(static_cast<char*>(temp) + l.hist_off);
    return sort_prepare(l, static_cast<char*>(temp), s);
}

hipError_t radix_sort_keys_prepared(uint64_t* keys[2], int32_t* sel, int32_t count, const uint32_t* count_dev,
                                    int32_t begin_bit, int32_t end_bit, void* temp, size_t temp_bytes, hipStream_t s,
                                    hipEvent_t passes_begin, hipEvent_t passes_end, const RangeArgs* ranges) {
    if (ranges && (ranges->tile_shift != 0 || ranges->tile_range == nullptr || begin_bit != 0))
        return hipErrorInvalidValue;
    if (count <= 1 || end_bit <= begin_bit) {
        if (!ranges) return hipSuccess;
        RangeArgs r = *ranges;   // no pass runs: the list is ranged as it stands
        r.keys = keys[*sel];
        r.size = count;
        r.size_dev = count_dev;
        return launch_ranges(r, s);
    }
    hipError_t e = sort_check<uint64_t>(count, begin_bit, end_bit, temp, temp_bytes);
    if (e != hipSuccess) return e;
    int32_t* none[2] = {nullptr, nullptr};
    return sort_passes_t<uint64_t>(keys, none, sel, count, count_dev, begin_bit, end_bit, static_cast<char*>(temp),
                                   1, s, passes_begin, passes_end, ranges ? ranges->tile_range : nullptr,
                                   ranges ? ranges->tiles : 0);
}

}  // namespace gr
