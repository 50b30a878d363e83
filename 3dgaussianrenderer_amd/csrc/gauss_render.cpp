// gauss_render.cpp -- the C++ object API (include/gauss_render.h) on top of
// the C-ABI.  Camera follows core/camera.cpp:7-55; Renderer follows
// core/renderer.cpp:39-139; GaussianModel::PreprocessSh implements what
// core/gaussian_model.cpp:9-11 leaves as a TODO.
#include "gauss_render.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>

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

CameraController::CameraController(Camera* camera) : camera_(camera) {}

void CameraController::Update(float delta_time) { (void)delta_time; }

void CameraController::OnMouseMove(float dx, float dy) {
    yaw_ += dx * sensitivity_;
    pitch_ += dy * sensitivity_;
    const float lim = 89.0f * (3.14159265358979f / 180.0f);
    if (pitch_ > lim) pitch_ = lim;
    if (pitch_ < -lim) pitch_ = -lim;
    UpdateCameraPose();
}

void CameraController::OnScroll(float dy) {
    distance_ -= dy * scroll_speed_ * distance_;
    distance_ = std::min(std::max(distance_, distance_ * 0.1f), distance_ * 10.0f);
    distance_ = std::min(std::max(distance_, 0.1f), 100.0f);
    UpdateCameraPose();
}

void CameraController::SetBounds(const vec3& mn, const vec3& mx) {
    bounds_min_ = mn;
    bounds_max_ = mx;
    center_ = vec3{(mn.x + mx.x) * 0.5f, (mn.y + mx.y) * 0.5f, (mn.z + mx.z) * 0.5f};
    const float dx = mx.x - mn.x, dy = mx.y - mn.y, dz = mx.z - mn.z;
    distance_ = std::sqrt((dx * dx + dy * dy) + dz * dz) * 0.8f;
    UpdateCameraPose();
}

void CameraController::SetOrbit(float yaw, float pitch, float distance) {
    const float lim = 89.0f * (3.14159265358979f / 180.0f);
    yaw_ = yaw;
    pitch_ = std::min(std::max(pitch, -lim), lim);
    distance_ = distance;
    UpdateCameraPose();
}

void CameraController::UpdateCameraPose() {
    float x = std::cos(yaw_) * std::cos(pitch_);
    float y = std::sin(pitch_);
    float z = std::sin(yaw_) * std::cos(pitch_);
    const float inv = 1.0f / std::sqrt((x * x + y * y) + z * z);
    x *= inv;
    y *= inv;
    z *= inv;
    camera_->SetPose(vec3{center_.x - x * distance_, center_.y - y * distance_, center_.z - z * distance_}, center_);
}

void FitCameraToBounds(Camera& camera, const vec3& mn, const vec3& mx) {
    const vec3 size{mx.x - mn.x, mx.y - mn.y, mx.z - mn.z};
    const vec3 center{mn.x + size.x * 0.5f, mn.y + size.y * 0.5f, mn.z + size.z * 0.5f};
    const float s = std::max(size.x, std::max(size.y, size.z));
    const float oy = s * 0.5f, oz = s * 0.5f;
    const float inv = 1.0f / std::sqrt(oy * oy + oz * oz);
    const float dist = std::sqrt(2.0f) * s * 0.5f / std::tan(camera.GetFovY() * 0.5f);
    camera.SetPose(vec3{center.x, center.y + oy * inv * dist, center.z + oz * inv * dist}, center);
}

bool WriteImage(const std::string& path, const uint8_t* rgba8, int width, int height) {
    return gr_image_write(path.c_str(), rgba8, width, height, 1) == GR_OK;
}

void StatsLog::AddFrame(const RenderStats& stats, double wall_ms) { frames_.push_back(Frame{stats, wall_ms}); }

bool StatsLog::WriteCsv(const std::string& path) const {
    std::FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "frame,sh_ms,project_ms,build_ms,sort_ms,raster_ms,total_ms,wall_ms,tile_entries\n");
    for (size_t i = 0; i < frames_.size(); ++i) {
        const RenderStats& s = frames_[i].s;
        std::fprintf(f, "%zu,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%lld\n", i, s.sh_eval_ms, s.project_ms,
                     s.build_tile_ms, s.sort_ms, s.raster_ms, s.total_ms, frames_[i].wall_ms,
                     (long long)s.tile_list_size);
    }
    return std::fclose(f) == 0;
}

namespace {
struct Agg {
    double mean, median, p99;
};
Agg aggregate(std::vector<double> v) {
    if (v.empty()) return Agg{0.0, 0.0, 0.0};
    double sum = 0.0;
    for (double x : v) sum += x;
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) {
        const double pos = q * (double)(v.size() - 1);
        const size_t lo = (size_t)pos;
        const size_t hi = std::min(lo + 1, v.size() - 1);
        return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
    };
    return Agg{sum / (double)v.size(), pct(0.5), pct(0.99)};
}
}  // namespace

std::string StatsLog::SummaryJson(int num_gaussians) const {
    const char* names[7] = {"sh_ms", "project_ms", "build_ms", "sort_ms", "raster_ms", "total_ms", "wall_ms"};
    std::string out = "{\"frames\": " + std::to_string(frames_.size()) +
                      ", \"gaussians\": " + std::to_string(num_gaussians);
    double k = 0.0;
    for (const Frame& fr : frames_) k += (double)fr.s.tile_list_size;
    char buf[256];
    for (int c = 0; c < 7; ++c) {
        std::vector<double> v;
        for (const Frame& fr : frames_) {
            const RenderStats& s = fr.s;
            const double x[7] = {s.sh_eval_ms, s.project_ms, s.build_tile_ms, s.sort_ms, s.raster_ms, s.total_ms,
                                 fr.wall_ms};
            v.push_back(x[c]);
        }
        const Agg a = aggregate(v);
        std::snprintf(buf, sizeof buf, ", \"%s\": {\"mean\": %.4f, \"median\": %.4f, \"p99\": %.4f}", names[c],
                      a.mean, a.median, a.p99);
        out += buf;
        if (c == 6) {
            std::snprintf(buf, sizeof buf, ", \"fps\": %.3f", a.mean > 0.0 ? 1000.0 / a.mean : 0.0);
            out += buf;
        }
    }
    std::snprintf(buf, sizeof buf, ", \"mean_tile_entries\": %.1f}",
                  frames_.empty() ? 0.0 : k / (double)frames_.size());
    return out + buf;
}

void StatsLog::PrintAverages(std::FILE* out) const {
    const double n = frames_.empty() ? 1.0 : (double)frames_.size();
    RenderStats sum;
    for (const Frame& fr : frames_) {
        sum.sh_eval_ms += fr.s.sh_eval_ms;
        sum.project_ms += fr.s.project_ms;
        sum.build_tile_ms += fr.s.build_tile_ms;
        sum.sort_ms += fr.s.sort_ms;
        sum.raster_ms += fr.s.raster_ms;
        sum.total_ms += fr.s.total_ms;
    }
    std::fprintf(out, "evaluateSphericalHarmonics average time ms: %2.6f\n", sum.sh_eval_ms / n);
    std::fprintf(out, "evaluateClipData average time ms: %2.6f\n", sum.project_ms / n);
    std::fprintf(out, "buildTileList average time ms: %2.6f\n", sum.build_tile_ms / n);
    std::fprintf(out, "sortTileList average time ms: %2.6f\n", sum.sort_ms / n);
    std::fprintf(out, "rasterize (ranges + tiles) average time ms: %2.6f\n", sum.raster_ms / n);
    std::fprintf(out, "Total average time ms: %2.6f\n", sum.total_ms / n);
}

int LoadPly(const std::string& path, GaussianModel& model) {
    int32_t n = 0, degree = 0, shc = 0;
    gr_status st = gr_ply_probe(path.c_str(), &n, &degree, &shc);
    if (st != GR_OK) {
        std::fprintf(stderr, "Failed to parse PLY: %s\n", gr_last_error());
        return 0;
    }
    const size_t nn = (size_t)std::max(n, 1);
    std::vector<float> pos(4 * nn), sr(4 * nn), col(4 * nn), rot(4 * nn);
    std::vector<float> sh(degree > 0 ? nn * (size_t)shc : 0);
    float mn[3], mx[3];
    int32_t got = 0;
    st = gr_ply_read(path.c_str(), n, pos.data(), sr.data(), col.data(), degree > 0 ? sh.data() : nullptr,
                     rot.data(), mn, mx, &got);
    if (st != GR_OK) {
        std::fprintf(stderr, "Failed to parse PLY: %s\n", gr_last_error());
        return 0;
    }
    model.gaussians.resize((size_t)got);
    for (int i = 0; i < got; ++i) {
        Gaussian& g = model.gaussians[i];
        std::memcpy(g.position, &pos[4 * (size_t)i], 12);
        g.opacity = pos[4 * (size_t)i + 3];
        std::memcpy(g.scale, &sr[4 * (size_t)i], 12);
        std::memcpy(g.rotation, &rot[4 * (size_t)i], 16);
        std::memcpy(g.color, &col[4 * (size_t)i], 12);
    }
    model.sh_coefficients = std::move(sh);
    model.sh_degree = degree;
    model.num_gaussians = got;
    model.bounds_min = vec3{mn[0], mn[1], mn[2]};
    model.bounds_max = vec3{mx[0], mx[1], mx[2]};
    model.PreprocessSh();
    return got;
}

bool SavePly(const std::string& path, const GaussianModel& model) {
    const int n = model.num_gaussians;
    if (n <= 0 || (int)model.gaussians.size() < n) return false;
    std::vector<float> pos((size_t)4 * n), sr((size_t)4 * n), col((size_t)4 * n);
    for (int i = 0; i < n; ++i) {
        const Gaussian& g = model.gaussians[i];
        std::memcpy(&pos[4 * (size_t)i], g.position, 12);
        pos[4 * (size_t)i + 3] = g.opacity;
        std::memcpy(&sr[4 * (size_t)i], g.scale, 12);
        const uint32_t q8 = gr_encode_vec4((g.rotation[1] + 1.0f) * 0.5f, (g.rotation[2] + 1.0f) * 0.5f,
                                           (g.rotation[3] + 1.0f) * 0.5f, (g.rotation[0] + 1.0f) * 0.5f);
        std::memcpy(&sr[4 * (size_t)i + 3], &q8, 4);
        std::memcpy(&col[4 * (size_t)i], g.color, 12);
        col[4 * (size_t)i + 3] = g.opacity;
    }
    return gr_ply_write(path.c_str(), n, pos.data(), sr.data(), col.data()) == GR_OK;
}

}  // namespace gauss_render
