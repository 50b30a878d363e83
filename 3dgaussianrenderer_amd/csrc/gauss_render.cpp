// gauss_render.cpp -- the C++ object API (include/gauss_render.h) on top of
// the C-ABI.  Camera follows core/camera.cpp:7-55; Renderer follows
// core/renderer.cpp:39-139; GaussianModel::PreprocessSh implements what
// core/gaussian_model.cpp:9-11 leaves as a TODO.
#include "gauss_render.h"

#include <cmath>
#include <cstring>
#include <limits>

namespace gauss_render {

namespace {
void check(gr_status s) {
    if (s != GR_OK) throw Error(s, gr_last_error());
}
}  // namespace

void GaussianModel::PreprocessSh() {
    const int shc = 3 * (sh_degree + 1) * (sh_degree + 1);
    if (sh_degree == 0 || num_gaussians == 0) {
        aligned_sh_.clear();
        return;
    }
    if (sh_coefficients.size() < (size_t)num_gaussians * shc)
        throw Error(GR_ERR_INVALID_ARGUMENT, "sh_coefficients holds fewer than N*3(d+1)^2 values");
    const size_t groups = ((size_t)num_gaussians + kDefaultShBlockSize - 1) / kDefaultShBlockSize;
    aligned_sh_.assign(groups * kDefaultShBlockSize * shc, 0.0f);
    check(gr_realign_spherical_harmonics(sh_coefficients.data(), aligned_sh_.data(), kDefaultShBlockSize, shc,
                                         num_gaussians));
}

void GaussianModel::UpdateBounds() {
    const float lo = std::numeric_limits<float>::lowest(), hi = std::numeric_limits<float>::max();
    bounds_min = vec3{hi, hi, hi};
    bounds_max = vec3{lo, lo, lo};
    for (const auto& g : gaussians) {
        bounds_min.x = std::min(bounds_min.x, g.position[0]);
        bounds_min.y = std::min(bounds_min.y, g.position[1]);
        bounds_min.z = std::min(bounds_min.z, g.position[2]);
        bounds_max.x = std::max(bounds_max.x, g.position[0]);
        bounds_max.y = std::max(bounds_max.y, g.position[1]);
        bounds_max.z = std::max(bounds_max.z, g.position[2]);
    }
}

Camera::Camera() { Update(); }

void Camera::SetFovY(float fov_y_rad) {
    fov_y_ = fov_y_rad;
    Update();
}
void Camera::SetResolution(int width, int height) {
    if (width <= 0 || height <= 0) throw Error(GR_ERR_INVALID_ARGUMENT, "bad resolution");
    width_ = width;
    height_ = height;
    Update();
}
void Camera::SetNearFar(float near_plane, float far_plane) {
    near_ = near_plane;
    far_ = far_plane;
    Update();
}
void Camera::SetPose(const vec3& position, const vec3& look_at) {
    position_ = position;
    look_at_ = look_at;
    Update();
}

void Camera::Update() {
    const float eye[3] = {position_.x, position_.y, position_.z};
    const float ctr[3] = {look_at_.x, look_at_.y, look_at_.z};
    check(gr_camera_make(eye, ctr, fov_y_, width_, height_, near_, far_, &data_));
}

mat4 Camera::GetViewProjection() const {
    mat4 m;
    std::memcpy(m.m, data_.view_projection, sizeof(m.m));
    return m;
}
mat4 Camera::GetProjection() const {
    mat4 m;
    std::memcpy(m.m, data_.projection, sizeof(m.m));
    return m;
}
mat4 Camera::GetView() const {
    mat4 m;
    std::memcpy(m.m, data_.view, sizeof(m.m));
    return m;
}
vec2 Camera::GetFovCotangent() const { return vec2{data_.fov_cotangent[0], data_.fov_cotangent[1]}; }
vec2 Camera::GetDepthScaleBias() const { return vec2{data_.depth_scale_bias[0], data_.depth_scale_bias[1]}; }
CameraData Camera::GetCameraData() const { return data_; }

Renderer::Renderer(int device, void* hip_stream) { check(gr_renderer_create(&r_, device, hip_stream)); }

Renderer::~Renderer() {
    if (r_) gr_renderer_destroy(r_);
}

void Renderer::UploadModel(const GaussianModel& model) {
    const int n = model.num_gaussians;
    if ((int)model.gaussians.size() < n) throw Error(GR_ERR_INVALID_ARGUMENT, "gaussians.size() < num_gaussians");
    std::vector<float> pos((size_t)4 * n), sr((size_t)4 * n), col((size_t)4 * n);
    for (int i = 0; i < n; ++i) {
        const Gaussian& g = model.gaussians[i];
        pos[4 * i + 0] = g.position[0];
        pos[4 * i + 1] = g.position[1];
        pos[4 * i + 2] = g.position[2];
        pos[4 * i + 3] = g.opacity;
        sr[4 * i + 0] = g.scale[0];
        sr[4 * i + 1] = g.scale[1];
        sr[4 * i + 2] = g.scale[2];
        // rotation is (w, x, y, z); the packed form is encodeVec4 of
        // ((x, y, z, w) + 1) / 2 (core/ply_parser.cpp:385)
        const uint32_t q8 = gr_encode_vec4((g.rotation[1] + 1.0f) * 0.5f, (g.rotation[2] + 1.0f) * 0.5f,
                                           (g.rotation[3] + 1.0f) * 0.5f, (g.rotation[0] + 1.0f) * 0.5f);
        std::memcpy(&sr[4 * i + 3], &q8, 4);
        col[4 * i + 0] = g.color[0];
        col[4 * i + 1] = g.color[1];
        col[4 * i + 2] = g.color[2];
        col[4 * i + 3] = g.opacity;
    }
    gr_model_host m{};
    m.num_gaussians = n;
    m.sh_degree = model.sh_degree;
    m.position = pos.data();
    m.scale_and_rotation = sr.data();
    m.color = col.data();
    m.sh = model.sh_degree > 0 ? model.sh_coefficients.data() : nullptr;
    check(gr_renderer_upload_model(r_, &m));
}

static void copy_stats(const gr_render_stats& s, RenderStats& out) {
    out.sh_eval_ms = s.sh_eval_ms;
    out.project_ms = s.project_ms;
    out.build_tile_ms = s.build_tile_ms;
    out.sort_ms = s.sort_ms;
    out.raster_ms = s.raster_ms;
    out.total_ms = s.total_ms;
    out.tile_list_size = s.tile_list_size;
}

void Renderer::Render(const Camera& camera, uint8_t* output_buffer, RenderStats& stats) {
    const CameraData cd = camera.GetCameraData();
    gr_render_stats s{};
    check(gr_renderer_render(r_, &cd, camera.GetWidth(), camera.GetHeight(), output_buffer, 1, &s));
    copy_stats(s, stats);
}

void Renderer::RenderToHost(const Camera& camera, uint8_t* host_rgba8, RenderStats& stats) {
    const CameraData cd = camera.GetCameraData();
    gr_render_stats s{};
    check(gr_renderer_render(r_, &cd, camera.GetWidth(), camera.GetHeight(), host_rgba8, 0, &s));
    copy_stats(s, stats);
}

}  // namespace gauss_render
