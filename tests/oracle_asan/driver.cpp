// driver.cpp -- runs one frame of the CPU oracle (oracle/mirror.cpp) from raw input files, for
// the sanitizer build of the oracle (tests/test_oracle_sanitizers.py).  Test infrastructure only.
//
//   driver <dir> <n> <degree> <W> <H>
// reads <dir>/position.f32 (n x 4), scale_rot.f32 (n x 4), sh.f32 (aligned SH), cam.f32 (23
// floats: view[16], fov cotangent[2], depth scale/bias[2], position[3]) and writes
// <dir>/image32.u8, keys.u64, values.i32 (the FP32 frame and its sorted tile list) and
// image64.u8 (the FP64 frame).  Prints "K=<entries> K64=<entries>".
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

extern "C" int64_t orc_render(const float* position, const float* scale_rot, const float* sh_aligned, int n,
                              int degree, const float* cam, int W, int H, float* color, float* xy, float* z,
                              float* ell, float* con, uint8_t* image, uint64_t* list_keys, int32_t* list_values,
                              int64_t list_capacity);
extern "C" int64_t orc64_render(const float* position, const float* scale_rot, const float* sh_aligned,
                                const float* color_in, int n, int degree, const float* cam, int W, int H,
                                uint8_t* image, double* image_float, uint64_t* list_keys, int32_t* list_values,
                                int64_t list_capacity);

template <typename T>
static std::vector<T> load(const std::string& path) {
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", path.c_str());
        std::exit(2);
    }
    std::fseek(f, 0, SEEK_END);
    const long bytes = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::vector<T> v((size_t)bytes / sizeof(T));
    if (std::fread(v.data(), sizeof(T), v.size(), f) != v.size()) std::exit(2);
    std::fclose(f);
    return v;
}

template <typename T>
static void save(const std::string& path, const T* p, size_t n) {
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(p, sizeof(T), n, f) != n) std::exit(3);
    std::fclose(f);
}

int main(int argc, char** argv) {
    if (argc != 6) return 1;
    const std::string dir = argv[1];
    const int n = std::atoi(argv[2]), degree = std::atoi(argv[3]), W = std::atoi(argv[4]), H = std::atoi(argv[5]);
    const auto pos = load<float>(dir + "/position.f32");
    const auto sr = load<float>(dir + "/scale_rot.f32");
    const auto sh = load<float>(dir + "/sh.f32");
    const auto cam = load<float>(dir + "/cam.f32");
    if ((int)pos.size() != 4 * n || (int)sr.size() != 4 * n || cam.size() != 23) return 4;
    std::vector<float> color((size_t)4 * n), xy((size_t)2 * n), z((size_t)n), ell((size_t)4 * n), con((size_t)4 * n);
    std::vector<uint8_t> img((size_t)W * H * 4), img64((size_t)W * H * 4);
    const int64_t k = orc_render(pos.data(), sr.data(), sh.data(), n, degree, cam.data(), W, H, color.data(),
                                 xy.data(), z.data(), ell.data(), con.data(), img.data(), nullptr, nullptr, 0);
    std::vector<uint64_t> keys((size_t)(k > 0 ? k : 1));
    std::vector<int32_t> vals((size_t)(k > 0 ? k : 1));
    orc_render(pos.data(), sr.data(), sh.data(), n, degree, cam.data(), W, H, color.data(), xy.data(), z.data(),
               ell.data(), con.data(), img.data(), keys.data(), vals.data(), k);
    const int64_t k64 = orc64_render(pos.data(), sr.data(), sh.data(), color.data(), n, degree, cam.data(), W, H,
                                     img64.data(), nullptr, nullptr, nullptr, 0);
    save(dir + "/image32.u8", img.data(), img.size());
    save(dir + "/keys.u64", keys.data(), (size_t)k);
    save(dir + "/values.i32", vals.data(), (size_t)k);
    save(dir + "/image64.u8", img64.data(), img64.size());
    std::printf("K=%lld K64=%lld\n", (long long)k, (long long)k64);
    return 0;
}
