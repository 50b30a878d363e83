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

}  // namespace gauss_render
