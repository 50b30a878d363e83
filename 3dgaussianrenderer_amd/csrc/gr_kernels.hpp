// gr_kernels.hpp -- kernel argument blocks and launcher declarations shared by
// the stage kernels (gr_stages.hip), the radix sort (gr_sort.hip) and the host
// side (gr_api.hip, gr_renderer.hip, gr_server.hip).  Arguments travel by
// value in the kernarg segment; there is no __constant__ state, so any number
// of contexts may run concurrently.
//
// Every stage kernel renders a batch of `nviews` views of one scene in one
// launch.  A single view (the reference's case) travels by value in `v0`; a
// batch reads its per-view descriptors from the device array `views`.  Per-view
// buffers are laid out view-major: element i of view v sits at v * count + i.
// Tile ids of a batch are global: view v owns tiles
// [views[v].tile_base, views[v].tile_base + views[v].tiles).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gr_math.hpp"

namespace gr {

constexpr int kShGroup = 256;   // Gaussians per SH group (app/main.cpp:278)

struct CameraArgs {
    m4 view;
    v2 fov_cotangent;
    v2 depth_scale_bias;
    v3 position;
    float aspect;
};

struct ViewDesc {
    CameraArgs cam;
    FrameGeom geom;
    int32_t tile_base;
    int32_t tiles;
    uchar4* out;            // RGBA8 image of this view (device)
};

__host__ __device__ __forceinline__ const ViewDesc& view_of(int v, const ViewDesc& v0, const ViewDesc* views) {
    return views ? views[v] : v0;
}

struct ShArgs {
    const float4* position;
    const void* sh;             // aligned [groups][sh_count][256]
    const uint2* packed;        // precision form: (scale x, y, z, opacity) 16-bit; else NULL
    float4* color;              // [nviews][count]
    int32_t count;
    int32_t sh_count;           // coefficients * 3
    int32_t nviews;
    ViewDesc v0;
    const ViewDesc* views;
};

struct ProjectArgs {
    const float4* position;
    const float4* scale_rotation;   // FP32 form
    const uint2* packed;            // precision form (scale + opacity, 16-bit)
    int32_t precision;              // 0 FP32, 1 FP16, 2 BF16
    float2* clip_xy;                // [nviews][count]
    float* clip_z;
    float4* ellipse;
    float4* conic;
    int32_t count;
    int32_t nviews;
    ViewDesc v0;
    const ViewDesc* views;
};

// The frame's front end in one pass (renderer and server): per Gaussian and view, SH colour,
// EWA projection and the visibility test (non-empty tile rect); each block of 256 Gaussians then
// writes its visible (view, Gaussian) pairs as build records, compacted in (view, index) order,
// into its own slice of a staging array (block b's records start at b * 256 * nviews) and their
// number to block_counts[b].  An exclusive scan of block_counts numbers the records globally:
// record r of block b is visible pair block_offsets[b] + r, so in (block, view, index) order --
// within one view, Gaussian-index order.  (A single pass that also placed the records globally
// would chain the blocks through a decoupled look-back, whose per-block ticket -- one atomic on
// one address -- costs 11.6 ns per block in series: 271 us for the 23,438 blocks of 6M
// Gaussians, measured.)
constexpr int kFrontThreads = 256;
constexpr int kFrontMaxViews = 64;
// Words to set before a kernel sequence runs: the front end sets them in passing (grid-stride
// over its lanes) so that the frame needs no separate fill or clear launches -- the tile ranges
// (-1 for tiles without entries) and a sort's zeroed scratch (radix_sort_zero_spans).
struct FillSpan {
    uint32_t* p;
    int64_t n;
    uint32_t value;
};
constexpr int kFrontFills = 5;

struct FrontArgs {
    FillSpan fills[kFrontFills];    // set before any other work of the frame reads them
    int32_t nfills;
    const float4* position;
    const float4* scale_rotation;   // FP32 form
    const uint2* packed;            // precision form (scale + opacity, 16-bit)
    const void* sh;                 // aligned [groups][sh_count][256] coefficients (degree > 0)
    int32_t degree;                 // 0: no SH pass (colour uploaded)
    int32_t sh_count;
    int32_t precision;              // 0 FP32, 1 FP16, 2 BF16
    float4* color;                  // [nviews][count] (degree > 0)
    float2* clip_xy;                // [nviews][count]
    float* clip_z;                  // [nviews][count]; NULL: not stored (the records hold the depth)
    float4* ellipse;                // [nviews][count]; NULL: not stored (the records hold it)
    float4* conic;
    float4* records;                // [blocks][256 * nviews][2] staged build records
    uint32_t* block_counts;         // [blocks] records of each block
    int32_t count;
    int32_t nviews;                 // 1 .. kFrontMaxViews
    ViewDesc v0;
    const ViewDesc* views;
};
inline int front_blocks(int32_t count) { return (count + kFrontThreads - 1) / kFrontThreads; }
hipError_t launch_front(const FrontArgs& a, hipStream_t stream);
// Record numbering after the scan: offsets = exclusive scan of block_counts over `blocks`
// blocks, *total their sum; writes slots[offsets[b] + r] = b * region + r for r < block_counts[b]
// (region = 256 * nviews, blocks * region < 2^32), so the build passes walk the records in
// full 64-record waves wherever they sit in the staging array.
hipError_t launch_record_slots(const uint32_t* offsets, const uint32_t* total, int32_t blocks, int32_t region,
                               uint32_t* slots, hipStream_t stream);

// Depth-ordered tile lists (the renderer's path): the visible Gaussians of
// every view are the front end's build records, numbered in (block, view,
// index) order.  The count pass walks the records in that order, one wave per
// 64 records (footprint sizes spread evenly over the waves), and writes each
// record's emit record and its (depth word, record index | entry count) pair;
// the pairs are sorted stably by depth word; a scan of the counts taken in
// depth order writes every record's output offset contiguously (prefix[q] for
// the q-th record in depth order) and the window table (the record holding
// every kEmitWindow-th entry); the emit pass walks the list in windows of
// kEmitWindow entries, one wave per window, reading the records' offsets and
// packed words contiguously and their emit records by record index, and
// writes the window's entries in order.  Entries of one tile therefore sit in
// (depth word, index) order -- the order the stable sort of (tile << 32 |
// depth) keys emitted in index order produces -- and only the tile bits remain
// to be sorted.  (An index-order emit, with the offsets scattered back to
// index order, measured slower: the scattered 4-byte offset stores.)
//
// Build record (32 bytes):
//   [0] centre x, centre y, cos, sin     [1] extent x, extent y,
//   depth word (bits), view * count + index (bits)
//
// Emit record (16 bytes, uint4), written by the count pass per record:
//   footprints of at most 64 candidate tiles:
//     x, y  mask of the passing candidates (candidate l = row-major index in the
//           footprint rect; bit l of x | y << 32)
//     z     view * count + index
//     w     local tile index of the rect's corner (y0 * tiles_x + x0, < 2^20)
//           | (rect width - 1) << 20
//   larger footprints:
//     x     (w & kEmitSide) the word of BuildArgs::side from which the count pass
//           listed the record's passing tiles (global tile ids, in entry order);
//           otherwise the staging slot of the build record, whose candidates the
//           emit then tests again
//     y     exact entry count, z view * count + index, w kEmitBig [| kEmitSide]
constexpr uint32_t kEmitBig = 1u << 31;
constexpr uint32_t kEmitSide = 1u << 30;
constexpr int kEmitCornerBits = 20;

struct BuildArgs {
    const float2* clip_xy;       // [nviews][count]
    const float* clip_z;
    const float4* ellipse;
    uint32_t* batch_counts;      // [nviews][batches]: counts (count pass) / exclusive offsets (emit pass)
    uint64_t* keys;              // index-order walk: (tile << 32 | depth word)
    uint64_t* entries;           // record walk: Gaussian index << 32 | tile id (placed in depth order)
    int32_t* values;             // Gaussian index within the scene
    const float4* records;       // record walk: the front end's staged records (NULL: the per-view
                                 // index-order walk over clip_xy/ellipse)
    const uint32_t* record_slots;// record walk: staging slot of record p (launch_record_slots)
    const uint32_t* record_total;// record walk: device-side number of records
    uint32_t* depth_keys;        // count pass: depth word of record p (depth sort keys)
    uint32_t* depth_words;       // count pass: record p | entry count << index_bits (depth sort values;
                                 // the count saturated at 2^(32 - index_bits) - 1, then the emit
                                 // record holds it)
    uint4* emit_records;         // count pass: emit record of record p
    uint32_t* side;              // count pass, one view: the passing tiles of footprints over 64
                                 // candidates, listed for the emit (NULL: the emit tests their
                                 // candidates again)
    int64_t side_words;          // capacity of side, cut into equal slices, one per count-pass wave
                                 // of the frame's records; a record that does not fit its wave's
                                 // slice is not listed
    int32_t index_bits;          // packed_index_bits(nviews * count)
    const uint32_t* sorted_words;   // emit pass: the packed record words in depth order
    const uint32_t* depth_prefix;   // emit pass: output offset of each depth-order position
    const uint32_t* windows;        // emit pass: depth-order position holding entry w * kEmitWindow
    const uint32_t* entry_total;    // emit pass: device-side number of entries (K)
    uint32_t* tile_hist;         // record emit: 256-bin histograms of the tile's digits (NULL: none)
    int32_t tile_passes;         // digits counted into tile_hist (<= 3)
    int32_t tile_digit_bits;     // their width: sort_digit_bits(0, tile bits)
    int32_t count;
    int32_t capacity;
    int32_t batches;             // ceil(count / 64) per view (index-order walk)
    int32_t nviews;
    ViewDesc v0;
    const ViewDesc* views;
};

struct RangeArgs {
    const uint64_t* keys;        // tile id in bits [tile_shift, tile_shift + 32)
    int32_t tile_shift;          // 32: (tile << 32 | depth word) keys; 0: (index << 32 | tile) entries
    int32_t* tile_range;
    const uint32_t* size_dev;    // device-side entry count (NULL: use `size`)
    int32_t size;                // entries, or the upper bound when size_dev is set
    int32_t tiles;
};

struct RasterArgs {
    const int32_t* tile_range;
    const int32_t* values;       // Gaussian index per sorted entry, or (when NULL)
    const uint64_t* entries;     // index << 32 | tile entries
    const float2* clip_xy;       // [nviews][count]
    const float4* conic;
    const float4* color;
    int32_t count;
    int32_t tiles;               // all views
    const uint32_t* order;       // tile of workgroup b (NULL: tile b)
    int32_t nviews;
    ViewDesc v0;
    const ViewDesc* views;
};

// Gaussians per build batch: one wave64 walks the candidate tiles of 64
// consecutive Gaussians.
constexpr int kBuildBatch = 64;
// Entries per window of the record emit (one wave writes one window of the list at a time).  At
// 6M Gaussians and 1080p: 128 / 256 / 512 / 1024 / 2048 entries, emit 181 / 154 / 146 / 144 / 150 us.
#ifndef GR_EMIT_WINDOW
#define GR_EMIT_WINDOW 512
#endif
constexpr int kEmitWindow = GR_EMIT_WINDOW;
// Words of the window table for a list of `capacity` entries.
inline int64_t emit_window_words(int64_t capacity) { return capacity / kEmitWindow + 1; }
constexpr int kBuildThreads = 256;
// Workgroups of the record emit pass (its waves take windows grid-stride, one tile-digit
// histogram flush per workgroup).
#ifndef GR_EMIT_BLOCKS
#define GR_EMIT_BLOCKS 4096
#endif
constexpr int kEmitBlocks = GR_EMIT_BLOCKS;

// launchers (all asynchronous on `stream`)
hipError_t launch_sh(const ShArgs& a, int degree, int precision, hipStream_t stream);
hipError_t launch_project(const ProjectArgs& a, hipStream_t stream);
hipError_t launch_build_count(const BuildArgs& a, hipStream_t stream);
// exclusive scan of n uint32 in place; scratch holds scan_scratch_words(n) words
int scan_scratch_words(int32_t n);
hipError_t launch_build_scan(uint32_t* data, int32_t n, uint32_t* total, uint32_t* scratch,
                             hipStream_t stream);
// Bits of the record index in a packed record word for n records (>= 1, <= 31).
int packed_index_bits(int64_t n);
// Exclusive scan of the counts held by packed record words (record index in
// bits [0, index_bits), count above it, a saturated count taken from the
// record's emit record): the first *valid words, in their order; the prefix
// of word j goes to prefix[j], the total to *total, and windows[w] (w <
// nwindows; nwindows 0: no table) receives the j whose entries [prefix[j],
// prefix[j] + count) hold entry w * kEmitWindow.
hipError_t launch_build_scan_packed(const uint32_t* words, int32_t n, const uint32_t* valid, const uint4* emit_records,
                                    int32_t index_bits, uint32_t* total, uint32_t* scratch, uint32_t* prefix,
                                    uint32_t* windows, int64_t nwindows, hipStream_t stream);
// The same window table from the scanned prefix of the first *valid of n records (K at *total),
// for a caller that sizes the table after the scan (nwindows 0 in launch_build_scan_packed).
hipError_t launch_emit_windows(const uint32_t* prefix, const uint32_t* valid, int32_t n, const uint32_t* total,
                               uint32_t* windows, int64_t nwindows, hipStream_t stream);
hipError_t launch_build_emit(const BuildArgs& a, hipStream_t stream);
// entries (index << 32 | tile) -> keys[i] = tile << 32 | depth word of clip_z[index], values[i] = index
hipError_t launch_expand_keys(const uint64_t* entries, const float* clip_z, uint64_t* keys, int32_t* values,
                              int64_t n, hipStream_t stream);
hipError_t launch_fill_i32(int32_t* p, int32_t value, int64_t n, hipStream_t stream);
hipError_t launch_ranges(const RangeArgs& a, hipStream_t stream);
hipError_t launch_raster(const RasterArgs& a, hipStream_t stream);
// order[0..tiles) = the tiles by decreasing entry count (classes of powers of two), for RasterArgs::order
hipError_t launch_tile_order(const int32_t* tile_range, int32_t tiles, uint32_t* order, hipStream_t stream);

// radix sort (gr_sort.hip)
size_t radix_sort_temp_bytes(int32_t count, int32_t begin_bit, int32_t end_bit);
// Digit width of a sort over bits [begin_bit, end_bit): ceil((end - begin) / 8) passes of at most
// 8 bits, as even as possible (13 tile bits: 7 + 6); pass p sorts bits
// [begin + p * w, min(begin + (p + 1) * w, end)).
int sort_digit_bits(int32_t begin_bit, int32_t end_bit);
// Sorts (keys[sel], values[sel]) on bits [begin_bit, end_bit); on return *sel
// names the buffer pair that holds the sorted data.  With values[0] == NULL
// the keys are sorted alone (the persistent keys-only passes).  The number of pairs is
// min(*count_dev, count) when count_dev is not NULL (a device-side count, so
// the sort can follow a build without a host round trip), else count; grids
// are sized for `count`.
hipError_t radix_sort_pairs(uint64_t* keys[2], int32_t* values[2], int32_t* sel, int32_t count,
                            const uint32_t* count_dev, int32_t begin_bit, int32_t end_bit, void* temp,
                            size_t temp_bytes, hipStream_t stream);
// Copies of the sort's digit histograms: a producer may add its counts to any
// copy (copy c starts kSortHistCopyStride words after copy 0); the scan sums them.
constexpr int kSortHistCopies = 16;
constexpr int kSortHistCopyStride = 8 * 256;
// Two-step form of the keys-only 64-bit sort for a producer that accumulates
// the digit histograms itself: radix_sort_prepare zeroes the histograms
// (passes x 256 uint32 at *hist, pass p counting digit (key >> (begin_bit + w p)) & (2^w - 1),
// w = sort_digit_bits(begin_bit, end_bit), in each of the kSortHistCopies copies), the partition tickets and the first
// look-back region of `temp` sized for `count`; after the producer has added
// exactly the keys the sort will see, radix_sort_keys_prepared scans them and
// runs the passes.  A device-side count above `count` then means the producer
// did not write the list: the prepared sort leaves the buffers untouched.
// Used for the renderer's tile-list entries (Gaussian index << 32 | tile id),
// sorted on the tile bits [0, end_bit).  ranges (optional; tile_shift 0, begin_bit 0): the
// sort also writes the tile ranges of the sorted list into ranges->tile_range for tiles
// [0, ranges->tiles), as launch_ranges would (its keys, size and size_dev are the sort's own);
// tiles without entries keep what they held, which must be -1 for both words.
hipError_t radix_sort_prepare(int32_t count, void* temp, size_t temp_bytes, uint32_t** hist, hipStream_t stream);
// The words radix_sort_prepare (and the pair sorts' own first step) zero, as two spans, for a
// caller that zeroes them in an earlier kernel of its own (FrontArgs fills); *hist as above.
hipError_t radix_sort_zero_spans(int32_t count, void* temp, size_t temp_bytes, FillSpan spans[2], uint32_t** hist);
hipError_t radix_sort_keys_prepared(uint64_t* keys[2], int32_t* sel, int32_t count, const uint32_t* count_dev,
                                    int32_t begin_bit, int32_t end_bit, void* temp, size_t temp_bytes,
                                    hipStream_t stream, hipEvent_t passes_begin = nullptr,
                                    hipEvent_t passes_end = nullptr, const RangeArgs* ranges = nullptr);
// The same for 32-bit keys (depth words); same temp layout and size.
// zeroed: the spans radix_sort_zero_spans names for `count` were zeroed by the caller after the
// scratch's last use (the sort then skips that step).
hipError_t radix_sort_pairs_u32(uint32_t* keys[2], int32_t* values[2], int32_t* sel, int32_t count,
                                const uint32_t* count_dev, int32_t begin_bit, int32_t end_bit, void* temp,
                                size_t temp_bytes, hipStream_t stream, bool zeroed = false);

}  // namespace gr
