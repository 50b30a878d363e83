/*
 * gr_capi.h -- C-ABI of the MI355X Gaussian-splat per-frame renderer.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, returns a
 * gr_status and never exits the process (the reference prints and calls exit:
 * core/cuda_error_check.h:2-11).  The message of the last failure on the
 * calling thread is returned by gr_last_error().
 *
 * Two levels, mirroring the two levels of the reference:
 *
 *  1. Stage level -- the launchers of core/gaussian_render.cuh:113-126
 *     (setGlobalArgs, setTileListArgs, evaluateSphericalHarmonics,
 *     evaluateSplatClipData, buildTileList, sortTileList, evaluateTileRange,
 *     rasterizeTile).  The reference keeps its arguments in `__constant__`
 *     globals (core/gaussian_render.cu:32-43), which makes it non-reentrant;
 *     here they live in a gr_context (one HIP stream each) and reach the
 *     kernels by value in the kernarg segment.  The caller owns every device
 *     buffer, exactly as app/main.cpp:295-313 does.
 *
 *  2. Object level -- gauss_render::Renderer::UploadModel / Render
 *     (core/renderer.h:13-50): gr_renderer_* own the device buffers and run
 *     the whole frame (core/renderer.cpp:94-139, app/main.cpp:326-433).
 *
 * Layouts (all device buffers, all FP32 unless stated):
 *   position            float4[N]  xyz world position, w activated opacity
 *   scale_and_rotation  float4[N]  xyz linear scale, w = bit pattern of the
 *                                  8-bit-per-component quaternion (x in bits
 *                                  24..31, then y, z, w), core/ply_parser.cpp:231-239
 *   color               float4[N]  rgb + opacity (SH output)
 *   spherical_harmonics float[ceil(N/256)][sh_count][256] ("aligned" layout,
 *                                  app/main.cpp:147-177); sh_count = 3(d+1)^2
 *   position_clip_xy    float2[N], position_clip_z float[N]
 *   screen_ellipse      float4[N]  (cos, sin, extent.x, extent.y)
 *   conic               float4[N]  (a, b, c, 0) of the inverse 2D covariance
 *   tile keys           uint64[K]  (tile << 32) | quantised depth
 *   tile values         int32[K]   Gaussian index
 *   tile_range          int32[2*T] [start, end) per tile, -1 when empty
 *   back_buffer         uchar4[W*H] RGBA8, row 0 = bottom (clip y = -1)
 */
#ifndef GR_CAPI_H_
#define GR_CAPI_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#define GR_EXTERN extern "C"
#else
#define GR_EXTERN extern
#endif
#define GR_API GR_EXTERN __attribute__((visibility("default")))

typedef enum gr_status {
    GR_OK = 0,
    GR_ERR_INVALID_ARGUMENT = 1,
    GR_ERR_HIP = 2,            /* a HIP runtime call failed */
    GR_ERR_OUT_OF_MEMORY = 3,
    GR_ERR_CAPACITY = 4,       /* tile list larger than the caller's capacity */
    GR_ERR_BUDGET = 5,         /* serving: request does not fit the HBM budget */
    GR_ERR_NOT_FOUND = 6
} gr_status;

/* SH storage precision of the "precision form" (FP32 accumulation always). */
typedef enum gr_precision {
    GR_PRECISION_FP32 = 0,
    GR_PRECISION_FP16 = 1,
    GR_PRECISION_BF16 = 2
} gr_precision;

/* core/gaussian_render.cuh:16-25 (CameraData).  Matrices are column-major,
 * element [c][r] at index 4*c + r, like glm::mat4. */
typedef struct gr_camera_data {
    float view_projection[16];
    float projection[16];
    float view[16];
    float fov_cotangent[2];
    float depth_scale_bias[2];
    float position[3];
    float aspect;
} gr_camera_data;

/* core/gaussian_render.cuh:56-72 (GlobalArgs), plus the output size that the
 * reference fixes at compile time (core/consts.h:4-8). */
typedef struct gr_global_args {
    gr_camera_data camera_data;
    const float* position;
    const float* scale_and_rotation;
    float* color;
    int32_t spherical_harmonics_degree;
    int32_t spherical_harmonics_count;     /* floats per Gaussian = 3(d+1)^2 */
    const void* spherical_harmonics;       /* aligned layout, see above */
    float* position_clip_xy;
    float* position_clip_z;
    float* conic;
    float* screen_ellipse;
    int32_t splat_count;
    uint8_t* back_buffer;
    int32_t* tile_range;
    int32_t width;                         /* pixels; reference: 1024 */
    int32_t height;                        /* pixels; reference: 1024 */
    int32_t sh_precision;                  /* gr_precision of spherical_harmonics */
    /* Precision form only (sh_precision != FP32): 4 x 16-bit per Gaussian
     * (scale x, y, z, opacity) in the same 16-bit format as the SH, and
     * position.w then holds the packed quaternion bits instead of opacity.
     * NULL in the FP32 form. */
    const void* packed_scale_opacity;
} gr_global_args;

/* core/gaussian_render.cuh:74-80 (TileListArgs). */
typedef struct gr_tile_list_args {
    uint64_t* keys;
    int32_t* values;
    int32_t size;
    int32_t capacity;
} gr_tile_list_args;

/* core/gaussian_render.cuh:82-111 (DoubleBuffer<T>). */
typedef struct gr_double_buffer_u64 {
    uint64_t* buffers[2];
    int32_t selector;
} gr_double_buffer_u64;

typedef struct gr_double_buffer_i32 {
    int32_t* buffers[2];
    int32_t selector;
} gr_double_buffer_i32;

/* core/types.h:26-33 (RenderStats), milliseconds; total_ms is end to end.
 * The renderer computes SH colour, projection and the compaction of the
 * visible Gaussians in one pass: sh_eval_ms is that pass and project_ms is 0. */
typedef struct gr_render_stats {
    double sh_eval_ms;
    double project_ms;
    double build_tile_ms;
    double sort_ms;
    double raster_ms;          /* includes the tile-range pass */
    double total_ms;
    int64_t tile_list_size;    /* K, tile-list entries of the frame */
} gr_render_stats;

/* Host-side scene in packed form (what core/ply_parser.cpp:241-393 produces). */
typedef struct gr_model_host {
    int32_t num_gaussians;
    int32_t sh_degree;                 /* 0..3 */
    const float* position;             /* float4[N] */
    const float* scale_and_rotation;   /* float4[N] */
    const float* color;                /* float4[N], used when sh_degree == 0; may be NULL otherwise */
    const float* sh;                   /* [N][3(d+1)^2], coefficient j channel c at 3j+c; NULL when d == 0 */
    int32_t precision;                 /* gr_precision of the device copy (FP32 = the reference's form) */
} gr_model_host;

typedef struct gr_context gr_context;   /* stream + stage arguments + build scratch */
typedef struct gr_timer gr_timer;       /* CudaTimer (core/utilities.h:155-187) */
typedef struct gr_renderer gr_renderer; /* object-level renderer */

/* ---- errors ---------------------------------------------------------- */
GR_API const char* gr_last_error(void);
GR_API const char* gr_version(void);

/* ---- context / timer (replace __constant__ globals and CudaTimer) ----- */
/* stream: a hipStream_t, or NULL to create (and own) a new one. */
GR_API gr_status gr_context_create(gr_context** out, void* stream);
/* Same, with flags: GR_CONTEXT_BORROW_STREAM uses `stream` as given even when
 * it is NULL (the legacy default stream). */
#define GR_CONTEXT_BORROW_STREAM 1
GR_API gr_status gr_context_create_ex(gr_context** out, void* stream, int32_t flags);
GR_API gr_status gr_context_destroy(gr_context* ctx);
GR_API gr_status gr_context_synchronize(gr_context* ctx);

/* CudaTimer: start/stop events on the context stream; elapsed_ms waits for
 * the stop event (core/utilities.h:173-182) -- the stage launchers never do.
 * The events are for timing only: they do not make device writes visible to
 * the host (use gr_context_synchronize or a copy for that). */
GR_API gr_status gr_timer_create(gr_timer** out);
GR_API gr_status gr_timer_destroy(gr_timer* timer);
GR_API gr_status gr_timer_elapsed_ms(gr_timer* timer, float* out_ms);

/* ---- stage level: core/gaussian_render.cuh:113-126 ------------------- */
/* setGlobalArgs: core/gaussian_render.cu:35-38 */
GR_API gr_status gr_set_global_args(gr_context* ctx, const gr_global_args* args);
/* setTileListArgs: core/gaussian_render.cu:40-43 */
GR_API gr_status gr_set_tile_list_args(gr_context* ctx, const gr_tile_list_args* args);
/* evaluateSphericalHarmonics: core/gaussian_render.cu:118-133 (timer may be NULL) */
GR_API gr_status gr_evaluate_spherical_harmonics(gr_context* ctx, gr_timer* timer, int32_t count);
/* evaluateSplatClipData: core/gaussian_render.cu:217-232 */
GR_API gr_status gr_evaluate_splat_clip_data(gr_context* ctx, gr_timer* timer, int32_t count);
/* buildTileList: core/gaussian_render.cu:566-589.  Deterministic: entries are
 * emitted in Gaussian-index order, tiles row-major within a Gaussian.  Writes
 * the first min(K, capacity) entries into the tile-list args' keys/values and
 * returns min(K, capacity) in *out_size and K in *out_required (may be NULL).
 * num_blocks is accepted for signature fidelity and ignored. */
GR_API gr_status gr_build_tile_list(gr_context* ctx, gr_timer* timer, int32_t num_blocks,
                                    int32_t tile_list_capacity, int32_t* out_size,
                                    int64_t* out_required);
/* sortTileList: core/gaussian_render.cu:591-632.  Stable LSD radix sort of
 * (key, value) pairs on bits [0, 32 + ceil(log2(tiles))).  temp storage is
 * grown with hipMalloc when too small (like the reference) and swapped double
 * buffers point at the sorted data on return. */
GR_API gr_status gr_sort_tile_list(gr_context* ctx, gr_timer* timer, int32_t tile_list_size,
                                   void** device_temp_storage, size_t* temp_storage_bytes,
                                   gr_double_buffer_u64* keys, gr_double_buffer_i32* values);
/* No counterpart in the reference: sortTileList for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same sort as gr_sort_tile_list, with
 * its scratch in a fixed 16 MiB device array of the library (one per GPU,
 * held until the process ends) instead of caller temp storage: no hipMalloc,
 * hipFree or other allocation, and no memset (its counters are cleared by a
 * fill kernel).  At most *max_pairs pairs (about 54M; written on every call,
 * also on failure, when max_pairs is not NULL); a larger list is refused with
 * GR_ERR_CAPACITY before anything is launched.  Every call on one GPU shares
 * that GPU's scratch: calls that overlap on one GPU (on different streams)
 * race; calls one after another on a stream do not. */
GR_API gr_status gr_sort_tile_list_static(gr_context* ctx, gr_timer* timer, int32_t tile_list_size,
                                          gr_double_buffer_u64* keys, gr_double_buffer_i32* values,
                                          int32_t* max_pairs);
/* Frees device memory that the library allocated for the caller (the sort
 * temp storage; the reference frees it with cudaFree, app/main.cpp:474-477). */
GR_API gr_status gr_device_free(void* device_ptr);
/* Generic form used by the tests: explicit bit range. */
GR_API gr_status gr_radix_sort_pairs(gr_context* ctx, int32_t count, int32_t begin_bit,
                                     int32_t end_bit, void** device_temp_storage,
                                     size_t* temp_storage_bytes, gr_double_buffer_u64* keys,
                                     gr_double_buffer_i32* values);
/* evaluateTileRange: core/gaussian_render.cu:662-677 (clears the ranges to -1
 * first, which app/main.cpp:365 does outside the launcher) */
GR_API gr_status gr_evaluate_tile_range(gr_context* ctx, gr_timer* timer, int32_t tile_list_size);
/* rasterizeTile: core/gaussian_render.cu:761-775.  Writes every pixel of the
 * back buffer (RGBA 0 for tiles without entries, app/main.cpp:363). */
GR_API gr_status gr_rasterize_tile(gr_context* ctx, gr_timer* timer);

/* ---- host helpers ------------------------------------------------------ */
/* realignSphericalHarmonics (app/main.cpp:147-177): AoS [N][sh_count] ->
 * [ceil(N/group)][sh_count][group].  dst holds ceil(N/group)*group*sh_count
 * floats; padding lanes are zero. */
GR_API gr_status gr_realign_spherical_harmonics(const float* src, float* dst, int32_t group_size,
                                                int32_t sh_count, int32_t splat_count);
/* Precision form (FP16/BF16 storage, FP32 arithmetic): converts an FP32
 * model into the device layout of that form, with round-to-nearest-even.
 *   position_out  float4[N]  (x, y, z, bit pattern of the packed quaternion)
 *   packed_out    uint16[4N] (scale x, y, z, opacity) as FP16 or BF16
 *   sh_out        uint16[ceil(N/256)*256*3(d+1)^2] aligned SH (may be NULL when d == 0) */
GR_API gr_status gr_pack_precision_model(const gr_model_host* model, int32_t precision, float* position_out,
                                         uint16_t* packed_out, uint16_t* sh_out);
/* Scalar conversions of the precision form (round to nearest even). */
GR_API uint16_t gr_float_to_half(float x);
GR_API uint16_t gr_float_to_bf16(float x);
/* encodeVec4 (core/ply_parser.cpp:231-239) */
GR_API uint32_t gr_encode_vec4(float x, float y, float z, float w);
/* Camera (core/camera.cpp:7-55): fill CameraData from a pose. */
GR_API gr_status gr_camera_make(const float eye[3], const float center[3], float fov_y_rad,
                                int32_t width, int32_t height, float z_near, float z_far,
                                gr_camera_data* out);
/* Number of tiles and sort bits for a W x H frame (16x16 tiles). */
GR_API gr_status gr_tile_grid(int32_t width, int32_t height, int32_t* tiles_x, int32_t* tiles_y,
                              int32_t* sort_end_bit);

/* ---- synthetic scenes (replace generateRandomGaussians, app/main.cpp:107-131) --
 * kind 1: "uniform cube" -- means U[-1,1]^3, log-scales N(ln 0.02, 0.3) per
 *         axis, Shoemake-uniform rotations, opacity sigmoid(N(0,1)).
 * kind 2: "clusters" -- 64 clusters, centres U[-5,5]^3, cluster covariance
 *         R diag(s^2) R^T with s log-uniform [0.1, 1.5]; per-Gaussian scales
 *         log-uniform [1e-3, 1e-1], rotations uniform, opacity sigmoid(N(0,1.5)).
 * SH (d > 0): DC ~ N(0.5, 0.2), higher bands ~ N(0, 0.05).  Each Gaussian's
 * stream is seeded from (seed, index), so the scene does not depend on the
 * number of host threads.  Arrays: position/scale_rot/color float4[N], sh
 * [N][3(d+1)^2] (may be NULL when d == 0).  centroid (optional) = mean of the
 * positions. */
typedef struct gr_scene_params {
    int32_t kind;
    int32_t sh_degree;
    int64_t count;
    uint64_t seed;
} gr_scene_params;
GR_API gr_status gr_generate_scene(const gr_scene_params* p, float* position, float* scale_and_rotation,
                                   float* color, float* sh, float centroid[3]);
/* Camera eye on a sphere of `radius` about `center`: pose `index` of a seeded
 * set (elevation within +-60 degrees, away from the lookAt singularity), or,
 * for orbit != 0, frame `index` of an orbit of `frames` frames at 20 degrees
 * elevation. */
GR_API gr_status gr_pose_eye(uint64_t seed, int32_t index, int32_t orbit, int32_t frames, const float center[3],
                             float radius, float eye[3]);

/* ---- PLY scenes: core/ply_parser.cpp (parseHeader :15-100, parsePly :102-131
 * and :241-393, savePly :143-219) ------------------------------------------
 * Binary little-endian PLY with float properties x y z rot_0..3 (rot_0 = w)
 * scale_0..2 f_dc_0..2 opacity [f_rest_0 ..]; the f_rest count fixes the SH
 * degree (0..3).  gr_ply_probe reads the header only: vertex count, degree
 * and sh_count = 3 + f_rest count.  gr_ply_read fills the packed model
 * arrays of parsePly: position float4[N] (x, y, z, sigmoid(opacity)),
 * scale_and_rotation float4[N] (exp(scale), encodeVec4 bits of the
 * normalised quaternion), color float4[N] (f_dc * SH_C0 + 0.5, opacity),
 * spherical_harmonics [N][sh_count] with coefficient j, channel c at 3j + c
 * (NULL allowed when the degree is 0), rotation_wxyz float4[N] the
 * normalised quaternion as floats (optional), bounds of the positions
 * (optional; the maximum starts at -FLT_MAX).  capacity is the number of
 * vertices the arrays hold.  Errors: GR_ERR_NOT_FOUND (cannot open),
 * GR_ERR_INVALID_ARGUMENT with the reference's message for a bad header,
 * missing property or wrong f_rest count, GR_ERR_CAPACITY.  gr_ply_write
 * writes savePly's 14-property file (rotation written in decodeVec4's
 * x, y, z, w order under rot_0..rot_3, as the reference does). */
GR_API gr_status gr_ply_probe(const char* path, int32_t* count, int32_t* sh_degree, int32_t* sh_count);
GR_API gr_status gr_ply_read(const char* path, int32_t capacity, float* position, float* scale_and_rotation,
                             float* color, float* spherical_harmonics, float* rotation_wxyz, float* bounds_min,
                             float* bounds_max, int32_t* count);
GR_API gr_status gr_ply_write(const char* path, int32_t count, const float* position,
                              const float* scale_and_rotation, const float* color);

/* ---- image output (replaces the GL display, app/main.cpp:435-466) -------
 * RGBA8 back buffer to a .png (RGBA, uncompressed deflate) or .ppm (RGB)
 * file chosen by the extension; bottom_up != 0 flips the rows so that row 0
 * of the buffer (the bottom of the image, clip y = -1) is written last. */
GR_API gr_status gr_image_write(const char* path, const uint8_t* rgba8, int32_t width, int32_t height,
                                int32_t bottom_up);

/* ---- object level: gauss_render::Renderer (core/renderer.h:13-50) ---- */
/* device < 0 keeps the current device; stream NULL creates one. */
GR_API gr_status gr_renderer_create(gr_renderer** out, int32_t device, void* stream);
GR_API gr_status gr_renderer_destroy(gr_renderer* r);
/* UploadModel: core/renderer.cpp:39-91 */
GR_API gr_status gr_renderer_upload_model(gr_renderer* r, const gr_model_host* model);
/* Render: core/renderer.cpp:94-139.  out_rgba8 holds width*height*4 bytes,
 * on the device when out_is_device != 0, else in host memory. */
GR_API gr_status gr_renderer_render(gr_renderer* r, const gr_camera_data* camera, int32_t width,
                                    int32_t height, uint8_t* out_rgba8, int32_t out_is_device,
                                    gr_render_stats* stats);
/* Device pointers of the last frame's intermediate buffers (for tests).
 * screen_ellipse is NULL: the renderer's front end keeps each visible
 * Gaussian's ellipse only in its build record.
 * sorted_keys/sorted_values is the frame's sorted tile list in the form of
 * the reference's sortTileList output (core/gaussian_render.cu:591-632):
 * (tile << 32 | depth word) keys.  The renderer keeps the list as tile ids
 * (the depth order is fixed before the tile sort), so this call makes the
 * full keys on the renderer's stream and waits for them. */
typedef struct gr_frame_view {
    const float* color;
    const float* position_clip_xy;
    const float* position_clip_z;
    const float* screen_ellipse;
    const float* conic;
    const uint64_t* sorted_keys;
    const int32_t* sorted_values;
    const int32_t* tile_range;
    int64_t tile_list_size;
    int32_t tiles_x;
    int32_t tiles_y;
    int32_t splat_count;
} gr_frame_view;
GR_API gr_status gr_renderer_frame_view(gr_renderer* r, gr_frame_view* out);

/* Timing of the renderer's frames (no counterpart in the reference, whose
 * CudaTimer synchronises after every stage, core/utilities.h:173-182).  HIP
 * events cost a few microseconds of idle GPU each, so a frame records only
 * what is asked for:
 *   stats != NULL in gr_renderer_render   the per-stage events of RenderStats
 *   GR_TIMING_SORT_PASSES                 two events around the tile-sort passes
 *   GR_TIMING_KERNELS                     an event after every kernel (or small group of
 *                                         kernels) of the frame, and the visible count
 *   GR_TIMING_FRONT                       two events around the front-end kernel (SH
 *                                         colour, projection and compaction), the frame's
 *                                         longest
 * A frame with stats == NULL at GR_TIMING_NONE (the default) records none. */
#define GR_TIMING_NONE 0
#define GR_TIMING_SORT_PASSES 1
#define GR_TIMING_KERNELS 2
#define GR_TIMING_FRONT 3
GR_API gr_status gr_renderer_set_timing(gr_renderer* r, int32_t level);

/* Per-kernel timing of the last frame, from HIP events on the renderer's
 * stream (for roofline reporting).  The tile-bit sort's onesweep passes run
 * back to back between two events: tile_sort_pass_ms is their total (-1 when
 * the frame did not time them), tile_sort_passes their number and
 * tile_list_size the pairs each pass moves.  At GR_TIMING_KERNELS,
 * kernel_ms[i] is the time of the work named kernel_name[i] (in frame order:
 * "front" (SH colour, projection and block-local compaction in one pass),
 * "numbering" (record numbers and staging slots), "record_count",
 * "depth_sort", "offsets", "emit", "tile_sort_scan", "tile_sort_passes"
 * (the last pass also writes the tile ranges), "raster") and
 * visible the number of Gaussians with a non-empty tile rect (else -1).  At
 * GR_TIMING_FRONT, front_ms is the front-end kernel's time (else -1). */
#define GR_KERNEL_TIMES_MAX 24
typedef struct gr_kernel_times {
    double tile_sort_pass_ms;
    int32_t tile_sort_passes;
    int32_t kernels;
    int64_t tile_list_size;
    int64_t visible;
    double kernel_ms[GR_KERNEL_TIMES_MAX];
    char kernel_name[GR_KERNEL_TIMES_MAX][24];
    double front_ms;
} gr_kernel_times;
GR_API gr_status gr_renderer_kernel_times(gr_renderer* r, gr_kernel_times* out);

/* ---- serving: continuous batching over resident scenes -----------------
 * No counterpart in the reference (one scene, one frame at a time,
 * app/main.cpp:326-472).  A server keeps scenes resident in HBM and renders
 * requests (scene, camera, width x height) submitted from any thread.  A
 * scheduler thread forms batches of pending requests of one scene and renders
 * each batch with ONE launch per stage (views fused: SH and projection read
 * the scene once for all views, the tile lists of all views go through one
 * sort), optionally replaying hipGraph-captured stage sequences (a batch
 * shape is captured when it repeats).  Images reach the callers' buffers
 * through copy-out threads, which post each completion after its copy.  Up to
 * `slots` batches are in flight on their own streams, so requests that
 * arrive while others run are admitted as soon as a slot and HBM budget are
 * free, and their work overlaps on the device.  Memory admission: resident
 * scenes + every slot's workspace (per-view buffers, tile keys, sort buffers)
 * stay within hbm_budget_bytes; a batch waits when its workspace would not
 * fit and fails with GR_ERR_BUDGET only when it cannot fit even alone.
 * A view renders bit-identically alone, in a batch, and through
 * gr_renderer_render (same precision). */
typedef struct gr_server gr_server;

typedef struct gr_server_config {
    int32_t device;               /* -1: current device */
    int32_t slots;                /* batches in flight (>= 1) */
    int32_t max_views_per_batch;  /* >= 1; a batch holds at most 64 views */
    int32_t use_graphs;           /* capture + replay repeated stage sequences with hipGraph */
    int64_t hbm_budget_bytes;     /* resident scenes + workspaces; 0 = no limit */
} gr_server_config;

typedef struct gr_request {
    int32_t scene;
    int32_t width;
    int32_t height;
    int32_t pad;
    gr_camera_data camera;
    uint8_t* out_rgba8;           /* host memory, width*height*4 bytes; written before completion */
    uint64_t user_tag;
} gr_request;

typedef struct gr_completion {
    uint64_t request_id;
    uint64_t user_tag;
    int32_t status;               /* gr_status */
    int32_t batch_views;          /* views fused in this request's batch */
    double submit_ms;             /* server clock (gr_server_clock_ms) */
    double start_ms;              /* batch launch */
    double finish_ms;             /* result in out_rgba8 */
    int64_t tile_list_size;       /* entries of the whole batch */
} gr_completion;

typedef struct gr_server_stats {
    int64_t resident_bytes;       /* scenes */
    int64_t workspace_bytes;      /* all slots */
    int64_t peak_bytes;
    int64_t batches;              /* batches completed */
    int64_t views;                /* views rendered in them */
    int64_t graph_launches;
    int64_t budget_waits;         /* times a batch waited for memory */
    int64_t graph_captures;       /* phase shapes captured as graphs (on their first repeat) */
    double capture_ms;            /* scheduler time spent capturing and instantiating graphs */
    double copy_out_ms;           /* copy-out thread time moving images into callers' buffers */
    int64_t workspace_reallocs;   /* slot buffers (re)allocated to fit larger batches */
    double realloc_ms;            /* scheduler time spent in those allocations */
} gr_server_stats;

GR_API gr_status gr_server_create(gr_server** out, const gr_server_config* config);
GR_API gr_status gr_server_destroy(gr_server* server);
/* Uploads a scene (precision field of the model selects FP32/FP16/BF16 storage). */
GR_API gr_status gr_server_add_scene(gr_server* server, const gr_model_host* model, int32_t* scene_id);
GR_API gr_status gr_server_remove_scene(gr_server* server, int32_t scene_id);
GR_API gr_status gr_server_submit(gr_server* server, const gr_request* request, uint64_t* request_id);
/* Returns up to max completions; waits up to wait_ms for the first one. */
GR_API gr_status gr_server_poll(gr_server* server, gr_completion* out, int32_t max, int32_t* n_out,
                                int32_t wait_ms);
GR_API gr_status gr_server_get_stats(gr_server* server, gr_server_stats* out);
GR_API double gr_server_clock_ms(gr_server* server);
/* Message of the last request that completed with an error status (empty if
 * none).  Requests fail on the server's scheduler thread, so their messages
 * never reach the submitting thread's gr_last_error(); this copies the latest
 * one (truncated to buf_bytes - 1 characters, NUL-terminated) into buf. */
GR_API gr_status gr_server_last_error(gr_server* server, char* buf, int32_t buf_bytes);

#endif /* GR_CAPI_H_ */
