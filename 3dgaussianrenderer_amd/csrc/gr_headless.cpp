// gr_headless.cpp -- headless host replacing the GLFW/OpenGL application
// (app/main.cpp).  It builds a synthetic scene, renders frames through the
// gauss_render::Renderer API and prints the per-stage averages that
// app/main.cpp:483-503 prints at exit, plus one JSON line.
//
//   gr_headless [--config 1|2|3] [--count N] [--width W] [--height H]
//               [--frames F] [--seed S] [--degree D] [--radius R]
//               [--fov DEG] [--orbit] [--out image.ppm]
//
// config 1: 10k uniform-cube Gaussians, 800x600, camera (0,0,3), fov 60.
// config 2: 1M clustered Gaussians, 1920x1080 orbit r=8, fov 50.
// config 3: 6M clustered Gaussians, 1920x1080, seeded poses r=8, fov 50.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gauss_render.h"

using namespace gauss_render;

namespace {

void write_ppm(const char* path, const std::vector<uint8_t>& rgba, int w, int h) {
    FILE* f = std::fopen(path, "wb");
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", path);
        return;
    }
    std::fprintf(f, "P6\n%d %d\n255\n", w, h);
    std::vector<uint8_t> row((size_t)w * 3);
    for (int y = h - 1; y >= 0; --y) {  // row 0 of the back buffer is the bottom
        for (int x = 0; x < w; ++x)
            for (int k = 0; k < 3; ++k) row[(size_t)3 * x + k] = rgba[((size_t)y * w + x) * 4 + k];
        std::fwrite(row.data(), 1, row.size(), f);
    }
    std::fclose(f);
}

}  // namespace

int main(int argc, char** argv) {
    int config = 1, width = -1, height = -1, frames = -1, degree = 3;
    long long count = -1;
    unsigned long long seed = 1;
    float radius = -1.0f, fov_deg = -1.0f;
    bool orbit = false;
    const char* out = nullptr;
    for (int i = 1; i < argc; ++i) {
        auto next = [&](const char* name) -> const char* {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "%s needs a value\n", name);
                std::exit(2);
            }
            return argv[++i];
        };
        if (!std::strcmp(argv[i], "--config")) config = std::atoi(next("--config"));
        else if (!std::strcmp(argv[i], "--count")) count = std::atoll(next("--count"));
        else if (!std::strcmp(argv[i], "--width")) width = std::atoi(next("--width"));
        else if (!std::strcmp(argv[i], "--height")) height = std::atoi(next("--height"));
        else if (!std::strcmp(argv[i], "--frames")) frames = std::atoi(next("--frames"));
        else if (!std::strcmp(argv[i], "--seed")) seed = std::strtoull(next("--seed"), nullptr, 10);
        else if (!std::strcmp(argv[i], "--degree")) degree = std::atoi(next("--degree"));
        else if (!std::strcmp(argv[i], "--radius")) radius = (float)std::atof(next("--radius"));
        else if (!std::strcmp(argv[i], "--fov")) fov_deg = (float)std::atof(next("--fov"));
        else if (!std::strcmp(argv[i], "--orbit")) orbit = true;
        else if (!std::strcmp(argv[i], "--out")) out = next("--out");
        else {
            std::fprintf(stderr, "unknown argument %s\n", argv[i]);
            return 2;
        }
    }
    const bool c1 = config == 1;
    if (count < 0) count = c1 ? 10000 : (config == 2 ? 1000000 : 6000000);
    if (width < 0) width = c1 ? 800 : 1920;
    if (height < 0) height = c1 ? 600 : 1080;
    if (frames < 0) frames = c1 ? 1 : (config == 2 ? 100 : 8);
    if (radius < 0) radius = c1 ? 3.0f : 8.0f;
    if (fov_deg < 0) fov_deg = c1 ? 60.0f : 50.0f;
    if (config == 2) orbit = true;

    try {
        gr_scene_params sp{c1 ? 1 : 2, degree, (int64_t)count, seed};
        const int shc = 3 * (degree + 1) * (degree + 1);
        std::vector<float> pos((size_t)count * 4), sr((size_t)count * 4), col((size_t)count * 4);
        std::vector<float> sh(degree > 0 ? (size_t)count * shc : 1);
        float centroid[3];
        gr_status st = gr_generate_scene(&sp, pos.data(), sr.data(), col.data(), degree > 0 ? sh.data() : nullptr,
                                         centroid);
        if (st != GR_OK) {
            std::fprintf(stderr, "scene generation failed\n");
            return 1;
        }
        if (c1) centroid[0] = centroid[1] = centroid[2] = 0.0f;

        gr_model_host m{(int32_t)count, degree, pos.data(), sr.data(), col.data(), degree > 0 ? sh.data() : nullptr};
        Renderer renderer;
        if (gr_renderer_upload_model(renderer.handle(), &m) != GR_OK) throw Error(GR_ERR_HIP, gr_last_error());

        Camera cam;
        cam.SetResolution(width, height);
        cam.SetFovY(fov_deg * 3.14159265358979f / 180.0f);
        cam.SetNearFar(0.01f, 100.0f);
        std::vector<uint8_t> image((size_t)width * height * 4);
        RenderStats sum{};
        double wall_ms = 0.0;
        for (int f = 0; f < frames; ++f) {
            float eye[3];
            if (c1) {
                eye[0] = 0.0f; eye[1] = 0.0f; eye[2] = radius;
            } else {
                gr_pose_eye(seed, f, orbit ? 1 : 0, frames, centroid, radius, eye);
            }
            cam.SetPose(vec3{eye[0], eye[1], eye[2]}, vec3{centroid[0], centroid[1], centroid[2]});
            RenderStats s;
            const auto t0 = std::chrono::steady_clock::now();
            renderer.RenderToHost(cam, image.data(), s);
            const auto t1 = std::chrono::steady_clock::now();
            wall_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
            sum.sh_eval_ms += s.sh_eval_ms;
            sum.project_ms += s.project_ms;
            sum.build_tile_ms += s.build_tile_ms;
            sum.sort_ms += s.sort_ms;
            sum.raster_ms += s.raster_ms;
            sum.total_ms += s.total_ms;
            sum.tile_list_size += s.tile_list_size;
        }
        const double n = frames > 0 ? (double)frames : 1.0;
        std::printf("evaluateSphericalHarmonics average time ms: %2.6f\n", sum.sh_eval_ms / n);
        std::printf("evaluateClipData average time ms: %2.6f\n", sum.project_ms / n);
        std::printf("buildTileList average time ms: %2.6f\n", sum.build_tile_ms / n);
        std::printf("sortTileList average time ms: %2.6f\n", sum.sort_ms / n);
        std::printf("rasterize (ranges + tiles) average time ms: %2.6f\n", sum.raster_ms / n);
        std::printf("Total average time ms: %2.6f\n", sum.total_ms / n);
        std::printf("{\"config\": %d, \"count\": %lld, \"width\": %d, \"height\": %d, \"frames\": %d, "
                    "\"sh_ms\": %.4f, \"project_ms\": %.4f, \"build_ms\": %.4f, \"sort_ms\": %.4f, "
                    "\"raster_ms\": %.4f, \"total_ms\": %.4f, \"wall_ms\": %.4f, \"mean_tile_entries\": %.1f}\n",
                    config, count, width, height, frames, sum.sh_eval_ms / n, sum.project_ms / n,
                    sum.build_tile_ms / n, sum.sort_ms / n, sum.raster_ms / n, sum.total_ms / n, wall_ms / n,
                    (double)sum.tile_list_size / n);
        if (out) write_ppm(out, image, width, height);
    } catch (const Error& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
