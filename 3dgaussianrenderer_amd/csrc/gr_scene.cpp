// gr_scene.cpp -- deterministic synthetic scenes and camera poses for the
// headless host (the reference's only generator is generateRandomGaussians,
// app/main.cpp:107-131, behind an `#if false`).
//
// Randomness: SplitMix64 (Steele, Lea & Flood 2014) seeded per Gaussian from
// (seed, index); normals by Box-Muller; uniform rotations by Shoemake's method
// (Graphics Gems III, 1992).  std::normal_distribution is avoided because its
// output is implementation-defined.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/gr_capi.h"

namespace {

struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double normal() {
        double u1 = uniform();
        const double u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

uint64_t mix_seed(uint64_t seed, uint64_t stream) {
    SplitMix m{seed ^ (stream * 0xd1b54a32d192ed03ull)};
    m.next();
    return m.next();
}

// Shoemake: uniform unit quaternion (x, y, z, w)
void shoemake(SplitMix& r, double q[4]) {
    const double u1 = r.uniform(), u2 = r.uniform(), u3 = r.uniform();
    const double a = std::sqrt(1.0 - u1), b = std::sqrt(u1);
    const double t1 = 6.283185307179586 * u2, t2 = 6.283185307179586 * u3;
    q[0] = a * std::sin(t1);
    q[1] = a * std::cos(t1);
    q[2] = b * std::sin(t2);
    q[3] = b * std::cos(t2);
}

void quat_to_mat(const double q[4], double R[3][3]) {  // row-major R[r][c]
    const double x = q[0], y = q[1], z = q[2], w = q[3];
    R[0][0] = 1 - 2 * (y * y + z * z); R[0][1] = 2 * (x * y - w * z);     R[0][2] = 2 * (x * z + w * y);
    R[1][0] = 2 * (x * y + w * z);     R[1][1] = 1 - 2 * (x * x + z * z); R[1][2] = 2 * (y * z - w * x);
    R[2][0] = 2 * (x * z - w * y);     R[2][1] = 2 * (y * z + w * x);     R[2][2] = 1 - 2 * (x * x + y * y);
}

float sigmoid(double x) { return (float)(1.0 / (1.0 + std::exp(-x))); }

struct Cluster {
    double c[3];
    double L[3][3];  // R * diag(s): x = c + L * n
};

template <typename F>
void parallel_for(int64_t n, F&& f) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
            for (int64_t i = lo; i < hi; ++i) f(i);
        });
    for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

gr_status gr_generate_scene(const gr_scene_params* p, float* position, float* scale_rot, float* color, float* sh,
                            float centroid[3]) {
    if (!p || !position || !scale_rot || !color || p->count < 0 || p->sh_degree < 0 || p->sh_degree > 3 ||
        (p->kind != 1 && p->kind != 2) || (p->sh_degree > 0 && !sh))
        return GR_ERR_INVALID_ARGUMENT;
    const int shc = 3 * (p->sh_degree + 1) * (p->sh_degree + 1);
    std::vector<Cluster> clusters;
    if (p->kind == 2) {
        SplitMix r{mix_seed(p->seed, 0xc105u)};
        clusters.resize(64);
        for (auto& cl : clusters) {
            for (int k = 0; k < 3; ++k) cl.c[k] = -5.0 + 10.0 * r.uniform();
            double q[4], R[3][3], s[3];
            shoemake(r, q);
            quat_to_mat(q, R);
            for (int k = 0; k < 3; ++k) s[k] = std::exp(std::log(0.1) + (std::log(1.5) - std::log(0.1)) * r.uniform());
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) cl.L[a][b] = R[a][b] * s[b];
        }
    }
    parallel_for(p->count, [&](int64_t i) {
        SplitMix r{mix_seed(p->seed, (uint64_t)i + 1)};
        double pos[3], sc[3], q[4], op;
        if (p->kind == 1) {
            for (int k = 0; k < 3; ++k) pos[k] = -1.0 + 2.0 * r.uniform();
            for (int k = 0; k < 3; ++k) sc[k] = std::exp(std::log(0.02) + 0.3 * r.normal());
            shoemake(r, q);
            op = r.normal();
        } else {
            const Cluster& cl = clusters[(size_t)(r.next() % 64u)];
            const double nrm[3] = {r.normal(), r.normal(), r.normal()};
            for (int a = 0; a < 3; ++a) pos[a] = cl.c[a] + cl.L[a][0] * nrm[0] + cl.L[a][1] * nrm[1] + cl.L[a][2] * nrm[2];
            for (int k = 0; k < 3; ++k) sc[k] = std::exp(std::log(1e-3) + (std::log(1e-1) - std::log(1e-3)) * r.uniform());
            shoemake(r, q);
            op = 1.5 * r.normal();
        }
        const float opacity = sigmoid(op);
        position[4 * i + 0] = (float)pos[0];
        position[4 * i + 1] = (float)pos[1];
        position[4 * i + 2] = (float)pos[2];
        position[4 * i + 3] = opacity;
        // encodeVec4(((x, y, z, w) + 1) * 0.5), core/ply_parser.cpp:385
        const uint32_t q8 = gr_encode_vec4(((float)q[0] + 1.0f) * 0.5f, ((float)q[1] + 1.0f) * 0.5f,
                                           ((float)q[2] + 1.0f) * 0.5f, ((float)q[3] + 1.0f) * 0.5f);
        float qf;
        std::memcpy(&qf, &q8, 4);
        scale_rot[4 * i + 0] = (float)sc[0];
        scale_rot[4 * i + 1] = (float)sc[1];
        scale_rot[4 * i + 2] = (float)sc[2];
        scale_rot[4 * i + 3] = qf;
        float dc[3];
        for (int k = 0; k < 3; ++k) dc[k] = (float)(0.5 + 0.2 * r.normal());
        // degree-0 colour of core/ply_parser.cpp:382-389 (f_dc * C0 + 0.5, unclamped)
        for (int k = 0; k < 3; ++k) color[4 * i + k] = dc[k] * 0.28209479177387814f + 0.5f;
        color[4 * i + 3] = opacity;
        if (p->sh_degree > 0) {
            float* s = sh + (size_t)i * shc;
            s[0] = dc[0];
            s[1] = dc[1];
            s[2] = dc[2];
            for (int k = 3; k < shc; ++k) s[k] = (float)(0.05 * r.normal());
        }
    });
    if (centroid) {
        double acc[3] = {0, 0, 0};
        for (int64_t i = 0; i < p->count; ++i)
            for (int k = 0; k < 3; ++k) acc[k] += position[4 * i + k];
        for (int k = 0; k < 3; ++k) centroid[k] = p->count ? (float)(acc[k] / (double)p->count) : 0.0f;
    }
    return GR_OK;
}

gr_status gr_pose_eye(uint64_t seed, int32_t index, int32_t orbit, int32_t frames, const float center[3],
                      float radius, float eye[3]) {
    if (!center || !eye || index < 0 || (orbit && frames <= 0)) return GR_ERR_INVALID_ARGUMENT;
    double az, el;
    if (orbit) {
        az = 6.283185307179586 * (double)index / (double)frames;
        el = 20.0 * 3.141592653589793 / 180.0;
    } else {
        SplitMix r{mix_seed(seed, 0x9053u + (uint64_t)index)};
        az = 6.283185307179586 * r.uniform();
        el = (-60.0 + 120.0 * r.uniform()) * 3.141592653589793 / 180.0;
    }
    eye[0] = center[0] + (float)(radius * std::cos(el) * std::sin(az));
    eye[1] = center[1] + (float)(radius * std::sin(el));
    eye[2] = center[2] + (float)(radius * std::cos(el) * std::cos(az));
    return GR_OK;
}

}  // extern "C"
