// gr_headless.cpp -- headless host replacing the GLFW/OpenGL application
// (app/main.cpp).  It loads a PLY scene (core/ply_parser.cpp, as app/main.cpp:262
// does) or builds a synthetic one, renders frames through the
// gauss_render::Renderer API along a scripted camera path, writes images and
// prints the per-stage averages that app/main.cpp:483-503 prints at exit,
// plus one JSON summary line.
//
//   gr_headless [--ply scene.ply | --config 1|2|3 [--count N] [--seed S] [--degree D]]
//               [--width W] [--height H] [--frames F] [--fov DEG]
//               [--camera poses|orbit|fit|controller] [--radius R] [--pitch DEG]
//               [--precision fp32|fp16|bf16]
//               [--out image.png|.ppm] [--out-pattern frames/f_%04d.png]
//               [--stats-csv frames.csv] [--stats-json summary.json]
//
// config 1: 10k uniform-cube Gaussians, 800x600, camera (0,0,3), fov 60.
// config 2: 1M clustered Gaussians, 1920x1080 orbit r=8, fov 50.
// config 3: 6M clustered Gaussians, 1920x1080, seeded poses r=8, fov 50.
// Camera paths (replacing the interactive controls, core/camera_controls.cpp and
// app/camera_controller.cpp):
//   poses       seeded eyes on a sphere of radius R about the scene centre
//   orbit       an orbit of F frames at 20 degrees elevation, radius R
//   fit         CameraControls::setBounds's viewpoint for the scene box
//   controller  CameraController: SetBounds, then yaw advancing 2 pi / F per
//               frame at --pitch (the mouse-drag orbit, scripted)
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gauss_render.h"

using namespace gauss_render;

int main(int argc, char** argv) {
    int config = 1, width = -1, height = -1, frames = -1, degree = 3, precision = GR_PRECISION_FP32;
    long long count = -1;
    unsigned long long seed = 1;
    float radius = -1.0f, fov_deg = -1.0f, pitch_deg = 20.0f;
    std::string camera_mode, ply, out, out_pattern, stats_csv, stats_json;
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
        else if (!std::strcmp(argv[i], "--pitch")) pitch_deg = (float)std::atof(next("--pitch"));
        else if (!std::strcmp(argv[i], "--camera")) camera_mode = next("--camera");
        else if (!std::strcmp(argv[i], "--orbit")) camera_mode = "orbit";
        else if (!std::strcmp(argv[i], "--ply")) ply = next("--ply");
        else if (!std::strcmp(argv[i], "--out")) out = next("--out");
        else if (!std::strcmp(argv[i], "--out-pattern")) out_pattern = next("--out-pattern");
        else if (!std::strcmp(argv[i], "--stats-csv")) stats_csv = next("--stats-csv");
        else if (!std::strcmp(argv[i], "--stats-json")) stats_json = next("--stats-json");
        else if (!std::strcmp(argv[i], "--precision")) {
            const std::string p = next("--precision");
            precision = p == "fp16" ? GR_PRECISION_FP16 : p == "bf16" ? GR_PRECISION_BF16 : GR_PRECISION_FP32;
            if (p != "fp16" && p != "bf16" && p != "fp32") {
                std::fprintf(stderr, "--precision must be fp32, fp16 or bf16\n");
                return 2;
            }
        } else {
            std::fprintf(stderr, "unknown argument %s\n", argv[i]);
            return 2;
        }
    }
    const bool from_ply = !ply.empty();
    const bool c1 = !from_ply && config == 1;
    if (count < 0) count = c1 ? 10000 : (config == 2 ? 1000000 : 6000000);
    if (width < 0) width = c1 || from_ply ? 800 : 1920;
    if (height < 0) height = c1 || from_ply ? 600 : 1080;
    if (frames < 0) frames = c1 ? 1 : (config == 2 ? 100 : 8);
    if (radius < 0) radius = c1 ? 3.0f : 8.0f;
    if (fov_deg < 0) fov_deg = c1 ? 60.0f : 50.0f;
    if (camera_mode.empty()) camera_mode = from_ply ? "fit" : (config == 2 ? "orbit" : (c1 ? "fixed" : "poses"));

    try {
        std::vector<float> pos, sr, col, sh;
        int n = 0, deg = 0;
        float center[3] = {0.0f, 0.0f, 0.0f};
        vec3 bmin{-1.0f, -1.0f, -1.0f}, bmax{1.0f, 1.0f, 1.0f};
        if (from_ply) {
            int32_t pn = 0, pd = 0, pshc = 0;
            if (gr_ply_probe(ply.c_str(), &pn, &pd, &pshc) != GR_OK) throw Error(GR_ERR_INVALID_ARGUMENT, gr_last_error());
            const size_t nn = (size_t)std::max(pn, 1);
            pos.resize(4 * nn);
            sr.resize(4 * nn);
            col.resize(4 * nn);
            sh.resize(pd > 0 ? nn * (size_t)pshc : 1);
            float mn[3], mx[3];
            int32_t got = 0;
            if (gr_ply_read(ply.c_str(), pn, pos.data(), sr.data(), col.data(), pd > 0 ? sh.data() : nullptr, nullptr,
                            mn, mx, &got) != GR_OK)
                throw Error(GR_ERR_INVALID_ARGUMENT, gr_last_error());
            n = got;
            deg = pd;
            bmin = vec3{mn[0], mn[1], mn[2]};
            bmax = vec3{mx[0], mx[1], mx[2]};
            for (int k = 0; k < 3; ++k) center[k] = 0.5f * (mn[k] + mx[k]);
            std::printf("Loaded %d gaussians (SH degree %d) from %s\n", n, deg, ply.c_str());
        } else {
            gr_scene_params sp{c1 ? 1 : 2, degree, (int64_t)count, seed};
            const int shc = 3 * (degree + 1) * (degree + 1);
            pos.resize((size_t)count * 4);
            sr.resize((size_t)count * 4);
            col.resize((size_t)count * 4);
            sh.resize(degree > 0 ? (size_t)count * shc : 1);
            if (gr_generate_scene(&sp, pos.data(), sr.data(), col.data(), degree > 0 ? sh.data() : nullptr, center) !=
                GR_OK)
                throw Error(GR_ERR_INVALID_ARGUMENT, gr_last_error());
            if (c1) center[0] = center[1] = center[2] = 0.0f;
            n = (int)count;
            deg = degree;
            bmin = vec3{center[0] - radius, center[1] - radius, center[2] - radius};
            bmax = vec3{center[0] + radius, center[1] + radius, center[2] + radius};
        }
        gr_model_host m{n, deg, pos.data(), sr.data(), col.data(), deg > 0 ? sh.data() : nullptr, precision};
        Renderer renderer;
        if (gr_renderer_upload_model(renderer.handle(), &m) != GR_OK) throw Error(GR_ERR_HIP, gr_last_error());

        Camera cam;
        cam.SetResolution(width, height);
        cam.SetFovY(fov_deg * 3.14159265358979f / 180.0f);
        cam.SetNearFar(0.01f, 100.0f);
        CameraController controller(&cam);
        if (camera_mode == "controller") controller.SetBounds(bmin, bmax);
        std::vector<uint8_t> image((size_t)width * height * 4);
        StatsLog log;
        for (int f = 0; f < frames; ++f) {
            const vec3 c{center[0], center[1], center[2]};
            if (camera_mode == "fixed") {
                cam.SetPose(vec3{0.0f, 0.0f, radius}, c);
            } else if (camera_mode == "fit") {
                FitCameraToBounds(cam, bmin, bmax);
            } else if (camera_mode == "controller") {
                controller.SetOrbit(6.28318530717959f * (float)f / (float)frames,
                                    pitch_deg * 3.14159265358979f / 180.0f, controller.distance());
            } else if (camera_mode == "orbit" || camera_mode == "poses") {
                float eye[3];
                gr_pose_eye(seed, f, camera_mode == "orbit" ? 1 : 0, frames, center, radius, eye);
                cam.SetPose(vec3{eye[0], eye[1], eye[2]}, c);
            } else {
                std::fprintf(stderr, "unknown camera mode %s\n", camera_mode.c_str());
                return 2;
            }
            RenderStats s;
            const auto t0 = std::chrono::steady_clock::now();
            renderer.RenderToHost(cam, image.data(), s);
            const auto t1 = std::chrono::steady_clock::now();
            log.AddFrame(s, std::chrono::duration<double, std::milli>(t1 - t0).count());
            if (!out_pattern.empty()) {
                char name[1024];
                std::snprintf(name, sizeof name, out_pattern.c_str(), f);
                if (!WriteImage(name, image.data(), width, height)) throw Error(GR_ERR_INVALID_ARGUMENT, gr_last_error());
            }
        }
        log.PrintAverages(stdout);
        std::printf("%s\n", log.SummaryJson(n).c_str());
        if (!out.empty() && !WriteImage(out, image.data(), width, height))
            throw Error(GR_ERR_INVALID_ARGUMENT, gr_last_error());
        if (!stats_csv.empty() && !log.WriteCsv(stats_csv)) throw Error(GR_ERR_INVALID_ARGUMENT, "cannot write " + stats_csv);
        if (!stats_json.empty()) {
            std::FILE* jf = std::fopen(stats_json.c_str(), "w");
            if (!jf) throw Error(GR_ERR_INVALID_ARGUMENT, "cannot write " + stats_json);
            std::fprintf(jf, "%s\n", log.SummaryJson(n).c_str());
            std::fclose(jf);
        }
    } catch (const Error& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
