// gauss_render.h -- C++ object API of the renderer, kept with the names and
// method signatures of the reference's gauss_render namespace:
//   Renderer        core/renderer.h:13-50   (UploadModel, Render)
//   GaussianModel   core/gaussian_model.h:21-41
//   Gaussian        core/gaussian_model.h:11-19
//   Camera          core/camera.h:11-59
//   RenderStats     core/types.h:26-33
//   CameraData / GlobalArgs / TileListArgs / DoubleBuffer
//                   core/gaussian_render.cuh:16-111 (here: the C-ABI structs)
// It sits on the C-ABI of gr_capi.h.  Unlike the reference (which prints and
// exits, core/cuda_error_check.h:2-11) failures throw gauss_render::Error.
#pragma once

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "gr_capi.h"

namespace gauss_render {

constexpr int kTileSize = 16;                 // core/types.h:10
constexpr int kDefaultScreenWidth = 1024;     // core/types.h:14
constexpr int kDefaultScreenHeight = 1024;    // core/types.h:16
constexpr int kDefaultShBlockSize = 256;      // core/types.h:18

using CameraData = gr_camera_data;
using GlobalArgs = gr_global_args;
using TileListArgs = gr_tile_list_args;

struct Error : std::runtime_error {
    gr_status status;
    Error(gr_status s, const std::string& what) : std::runtime_error(what), status(s) {}
};

struct vec3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};
struct vec2 {
    float x = 0.0f, y = 0.0f;
};
struct mat4 {
    float m[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};  // column-major
};

// core/types.h:26-33; total_ms is end to end (the reference app's printed
// total leaves SH out, app/main.cpp:490-495).
struct RenderStats {
    double sh_eval_ms = 0.0;
    double project_ms = 0.0;
    double build_tile_ms = 0.0;
    double sort_ms = 0.0;
    double raster_ms = 0.0;
    double total_ms = 0.0;
    int64_t tile_list_size = 0;
};

// core/gaussian_model.h:11-19.  Values are activated as the PLY loader leaves
// them (core/ply_loader.cpp:218-246): opacity in [0,1], linear scale,
// normalised rotation (w, x, y, z), colour = f_dc * C0 + 0.5.
struct Gaussian {
    float position[3];
    float opacity;
    float scale[3];
    float rotation[4];
    float color[3];
};

// core/gaussian_model.h:21-41.  sh_coefficients is [N][3(d+1)^2] with
// coefficient j, channel c at 3j + c (core/ply_parser.cpp:311-349).
class GaussianModel {
public:
    std::vector<Gaussian> gaussians;
    std::vector<float> sh_coefficients;
    int sh_degree = 0;
    int num_gaussians = 0;
    vec3 bounds_min;
    vec3 bounds_max;

    // Builds the aligned [ceil(N/256)][3(d+1)^2][256] layout read by the SH
    // kernel (the reference leaves this a stub, core/gaussian_model.cpp:9-11;
    // app/main.cpp:147-177 is the working version).
    void PreprocessSh();
    const std::vector<float>& GetAlignedSh() const { return aligned_sh_; }
    vec3 GetBoundsMin() const { return bounds_min; }
    vec3 GetBoundsMax() const { return bounds_max; }
    // Recomputes the bounds from the positions (lowest(), not min(): B19 of the
    // reference's parser, core/ply_parser.cpp:352).
    void UpdateBounds();

private:
    std::vector<float> aligned_sh_;
};

// core/camera.h:11-59, core/camera.cpp:7-55.
class Camera {
public:
    Camera();
    void SetFovY(float fov_y_rad);
    void SetResolution(int width, int height);
    void SetNearFar(float near_plane, float far_plane);
    void SetPose(const vec3& position, const vec3& look_at);

    mat4 GetViewProjection() const;
    mat4 GetProjection() const;
    mat4 GetView() const;
    vec3 GetPosition() const { return position_; }
    float GetAspect() const { return (float)width_ / (float)height_; }
    vec2 GetFovCotangent() const;
    vec2 GetDepthScaleBias() const;
    int GetWidth() const { return width_; }
    int GetHeight() const { return height_; }
    float GetFovY() const { return fov_y_; }
    float GetZNear() const { return near_; }
    float GetZFar() const { return far_; }
    // The CameraData block handed to the kernels (app/main.cpp:346-358).
    CameraData GetCameraData() const;

private:
    void Update();
    vec3 position_{0.0f, 0.0f, 3.0f};
    vec3 look_at_{0.0f, 0.0f, 0.0f};
    float fov_y_ = 0.872664626f;  // 50 degrees (core/camera.h:53)
    float near_ = 0.01f;
    float far_ = 100.0f;
    int width_ = kDefaultScreenWidth;
    int height_ = kDefaultScreenHeight;
    CameraData data_{};
};

// core/renderer.h:13-50.
class Renderer {
public:
    explicit Renderer(int device = -1, void* hip_stream = nullptr);
    ~Renderer();
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;

    void UploadModel(const GaussianModel& model);
    // output_buffer is a device RGBA8 buffer of W*H*4 bytes (the reference
    // renders into a mapped GL PBO, app/main.cpp:361,380).
    void Render(const Camera& camera, uint8_t* output_buffer, RenderStats& stats);
    // Same frame, copied to host memory.
    void RenderToHost(const Camera& camera, uint8_t* host_rgba8, RenderStats& stats);
    gr_renderer* handle() { return r_; }

private:
    gr_renderer* r_ = nullptr;
};

// app/camera_controller.h:12-40 / app/camera_controller.cpp: orbit about the
// scene centre.  yaw/pitch/distance drive the pose: direction = normalize(
// (cos yaw cos pitch, sin pitch, sin yaw cos pitch)), position = centre -
// direction * distance; pitch is clamped to +-89 degrees; SetBounds centres on
// the box and sets distance = 0.8 |max - min|.  SetOrbit is the headless
// driver (an orbit path sets yaw per frame instead of mouse deltas).
class CameraController {
public:
    explicit CameraController(Camera* camera);
    void Update(float delta_time);
    void OnMouseMove(float dx, float dy);
    void OnScroll(float dy);
    void SetBounds(const vec3& min_bounds, const vec3& max_bounds);
    void SetOrbit(float yaw, float pitch, float distance);
    float yaw() const { return yaw_; }
    float pitch() const { return pitch_; }
    float distance() const { return distance_; }
    vec3 center() const { return center_; }

private:
    void UpdateCameraPose();
    Camera* camera_;
    float yaw_ = 0.0f;
    float pitch_ = 0.0f;
    float distance_ = 2.0f;
    vec3 center_{};
    vec3 bounds_min_{-5.0f, -5.0f, -5.0f};
    vec3 bounds_max_{5.0f, 5.0f, 5.0f};
    float sensitivity_ = 0.01f;
    float scroll_speed_ = 0.1f;
};

// core/camera_controls.cpp:130-143 (CameraControls::setBounds): the default
// viewpoint for a scene box -- centre + normalize(0, s/2, s/2) * dist with
// s the largest box side and dist = sqrt(2) * s/2 / tan(fov_y/2), looking at
// the centre.
void FitCameraToBounds(Camera& camera, const vec3& min_bounds, const vec3& max_bounds);

// Back buffer (RGBA8, row 0 at the bottom) to a .png or .ppm file, top row
// first (gr_image_write).  Replaces the GL display, app/main.cpp:435-466.
bool WriteImage(const std::string& path, const uint8_t* rgba8, int width, int height);

// Per-frame RenderStats reporting (app/main.cpp:483-503 prints the averages;
// app/perf_overlay.cpp:115-161 draws FPS and the stage times over the image):
// one CSV row per frame, a JSON summary (mean, median and p99 of each stage
// and of the wall time per frame, FPS), and the reference's printed averages.
class StatsLog {
public:
    void AddFrame(const RenderStats& stats, double wall_ms);
    size_t frames() const { return frames_.size(); }
    bool WriteCsv(const std::string& path) const;
    std::string SummaryJson(int num_gaussians) const;
    void PrintAverages(std::FILE* out) const;

private:
    struct Frame {
        RenderStats s;
        double wall_ms;
    };
    std::vector<Frame> frames_;
};

// core/ply_loader.h:13 (LoadPly), with the parsing rules of
// core/ply_parser.cpp (gr_ply_read): fills gaussians (activated values,
// rotation (w, x, y, z) normalised), sh_coefficients ([N][3(d+1)^2], 3j + c),
// sh_degree, num_gaussians and the bounds, then PreprocessSh().  Returns the
// Gaussian count, or 0 after printing the reason (as the reference does); the
// reason also stays in gr_last_error().
int LoadPly(const std::string& path, GaussianModel& model);

// savePly (core/ply_parser.cpp:143-219) of the model's Gaussians; false on error.
bool SavePly(const std::string& path, const GaussianModel& model);

}  // namespace gauss_render
