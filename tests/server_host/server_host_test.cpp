// server_host_test.cpp -- drives the serving scheduler (gr_server_*) against
// the mocked runtime in mock_hip.cpp.  Test infrastructure only.
//
//   server_host_test mix <graphs> <budget_bytes>   alone + fused rounds over two scenes
//   server_host_test add_budget                    a scene larger than the budget is refused
//   server_host_test add_exact                     a scene within 1.1x of the budget is admitted exactly
//   server_host_test add_fail                      an allocation failing inside add_scene leaks nothing
//   server_host_test threads <graphs> [budget]     concurrent submitters, a scene removed mid-run
//
// Each request's camera position.x carries a tag in [1, 250]; the mocked
// raster paints that view with it, so every completed image is checked to be
// its own request's.  Prints one "RESULT key=value ..." line.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gr_capi.h"

extern "C" void mock_fail_malloc_after(int k);   // mock_hip.cpp

namespace {

struct Model {
    std::vector<float> pos, sr, col, sh;
    gr_model_host host{};
    Model(int n, int degree, int precision) {
        pos.assign((size_t)n * 4, 0.5f);
        sr.assign((size_t)n * 4, 0.1f);
        col.assign((size_t)n * 4, 0.2f);
        sh.assign((size_t)n * 3 * (degree + 1) * (degree + 1), 0.3f);
        host = gr_model_host{n, degree, pos.data(), sr.data(), col.data(), degree ? sh.data() : nullptr, precision};
    }
};

struct Req {
    std::vector<uint8_t> img;
    int tag = 0;
    bool no_output = false;   // submitted without an output buffer
};

struct Tally {
    std::mutex mu;
    std::unordered_map<uint64_t, Req*> live;
    long ok = 0, budget = 0, other = 0, bad_image = 0, max_batch = 0;
};

const int kRes[5][2] = {{640, 360}, {1280, 720}, {800, 600}, {333, 257}, {1920, 1080}};

uint64_t submit(gr_server* s, int32_t scene, int w, int h, int tag, Tally& t, std::vector<Req*>& own,
                bool no_output = false) {
    Req* r = new Req();
    r->img.assign((size_t)w * h * 4, 0);
    r->no_output = no_output;
    r->tag = tag;
    gr_request q{};
    q.scene = scene;
    q.width = w;
    q.height = h;
    q.camera.position[0] = (float)tag;
    q.out_rgba8 = no_output ? nullptr : r->img.data();
    q.user_tag = (uint64_t)tag;
    uint64_t id = 0;
    std::lock_guard<std::mutex> lk(t.mu);   // register before the completion can be polled
    if (gr_server_submit(s, &q, &id) != GR_OK) {
        delete r;
        return 0;
    }
    t.live[id] = r;
    own.push_back(r);
    return id;
}

int drain(gr_server* s, Tally& t, int want) {
    int got = 0;
    while (got < want) {
        gr_completion c[32];
        int32_t n = 0;
        if (gr_server_poll(s, c, 32, &n, 200) != GR_OK) return -1;
        std::lock_guard<std::mutex> lk(t.mu);
        for (int k = 0; k < n; ++k) {
            Req* r = t.live.at(c[k].request_id);
            t.live.erase(c[k].request_id);
            if ((int)c[k].user_tag != r->tag) ++t.bad_image;
            if (c[k].status == GR_OK) {
                ++t.ok;
                // an image equal to the request's tag; a request without an output buffer
                // leaves its (zeroed) buffer alone
                const uint8_t want = r->no_output ? 0 : (uint8_t)r->tag;
                for (uint8_t b : r->img)
                    if (b != want) {
                        ++t.bad_image;
                        break;
                    }
            } else if (c[k].status == GR_ERR_BUDGET) {
                ++t.budget;
            } else {
                ++t.other;
            }
            if (c[k].batch_views > t.max_batch) t.max_batch = c[k].batch_views;
        }
        got += n;
    }
    return got;
}

void print_result(gr_server* s, Tally& t, const char* extra = "") {
    gr_server_stats st{};
    gr_server_get_stats(s, &st);
    printf("RESULT ok=%ld budget=%ld other=%ld bad_image=%ld max_batch=%ld batches=%lld views=%lld graphs=%lld "
           "peak=%lld waits=%lld resident=%lld %s\n",
           t.ok, t.budget, t.other, t.bad_image, t.max_batch, (long long)st.batches, (long long)st.views,
           (long long)st.graph_launches, (long long)st.peak_bytes, (long long)st.budget_waits,
           (long long)st.resident_bytes, extra);
    fflush(stdout);
}

int run_mix(int graphs, long long budget) {
    gr_server_config cfg{-1, 3, 8, graphs, budget};
    gr_server* s = nullptr;
    if (gr_server_create(&s, &cfg) != GR_OK) return 2;
    Model m1(80000, 3, 0), m2(10000, 3, 1);
    int32_t a = -1, b = -1;
    if (gr_server_add_scene(s, &m1.host, &a) != GR_OK || gr_server_add_scene(s, &m2.host, &b) != GR_OK) return 3;
    Tally t;
    std::vector<Req*> own;
    int tag = 1;
    for (int rnd = 0; rnd < 3; ++rnd) {
        for (int i = 0; i < 2; ++i) {   // alone
            submit(s, a, kRes[i][0], kRes[i][1], tag, t, own);
            tag = tag % 250 + 1;
            drain(s, t, 1);
        }
        int want = 0;
        for (int i = 0; i < 6; ++i, ++want) {
            submit(s, a, kRes[i % 4][0], kRes[i % 4][1], tag, t, own, i == 4);   // one without output
            tag = tag % 250 + 1;
        }
        for (int i = 0; i < 3; ++i, ++want) {
            submit(s, b, kRes[i % 4][0], kRes[i % 4][1], tag, t, own);
            tag = tag % 250 + 1;
        }
        drain(s, t, want);
    }
    print_result(s, t);
    if (gr_server_remove_scene(s, b) != GR_OK) return 4;
    gr_server_destroy(s);
    for (Req* r : own) delete r;
    return 0;
}

int run_add_budget() {
    gr_server_config cfg{-1, 2, 4, 0, 1000000};
    gr_server* s = nullptr;
    if (gr_server_create(&s, &cfg) != GR_OK) return 2;
    Model m(80000, 3, 0);
    int32_t id = -1;
    const gr_status st = gr_server_add_scene(s, &m.host, &id);
    printf("RESULT add_status=%d\n", (int)st);
    gr_server_destroy(s);
    return 0;
}

// A scene whose exact size fits the budget with less than the workspaces' 25 % headroom to
// spare is admitted and accounted at its exact size.
int run_add_exact() {
    Model m(80000, 3, 0);
    gr_server_config cfg{-1, 2, 4, 0, 0};
    gr_server* s = nullptr;
    if (gr_server_create(&s, &cfg) != GR_OK) return 2;
    int32_t id = -1;
    if (gr_server_add_scene(s, &m.host, &id) != GR_OK) return 3;
    gr_server_stats st{};
    gr_server_get_stats(s, &st);
    const long long exact = st.resident_bytes;
    gr_server_destroy(s);
    const long long budget = exact + exact / 10;   // 1.1x the scene
    gr_server_config c2{-1, 2, 4, 0, budget};
    if (gr_server_create(&s, &c2) != GR_OK) return 2;
    const gr_status add = gr_server_add_scene(s, &m.host, &id);
    gr_server_get_stats(s, &st);
    printf("RESULT add_status=%d exact=%lld resident=%lld peak=%lld budget=%lld\n", (int)add, exact,
           (long long)st.resident_bytes, (long long)st.peak_bytes, budget);
    gr_server_destroy(s);
    return 0;
}

// An allocation that fails partway through gr_server_add_scene frees the buffers allocated
// before it: nothing stays counted as resident (and LeakSanitizer sees no leak at exit).
int run_add_fail() {
    Model m(80000, 3, 0);
    gr_server_config cfg{-1, 2, 4, 0, 0};
    gr_server* s = nullptr;
    if (gr_server_create(&s, &cfg) != GR_OK) return 2;
    int32_t id = -1;
    mock_fail_malloc_after(2);   // position and scale/rotation succeed, the SH buffer fails
    const gr_status add = gr_server_add_scene(s, &m.host, &id);
    mock_fail_malloc_after(-1);
    gr_server_stats st{};
    gr_server_get_stats(s, &st);
    const long long after_fail = st.resident_bytes;
    const gr_status again = gr_server_add_scene(s, &m.host, &id);   // the server still admits it
    gr_server_get_stats(s, &st);
    printf("RESULT add_status=%d resident_after_fail=%lld retry_status=%d resident=%lld\n", (int)add,
           after_fail, (int)again, (long long)st.resident_bytes);
    gr_server_destroy(s);
    return 0;
}

int run_threads(int graphs, long long budget) {
    gr_server_config cfg{-1, 3, 16, graphs, budget};
    gr_server* s = nullptr;
    if (gr_server_create(&s, &cfg) != GR_OK) return 2;
    Model m1(50000, 3, 0), m2(20000, 2, 2), m3(30000, 0, 0);
    int32_t ids[3];
    if (gr_server_add_scene(s, &m1.host, &ids[0]) != GR_OK || gr_server_add_scene(s, &m2.host, &ids[1]) != GR_OK ||
        gr_server_add_scene(s, &m3.host, &ids[2]) != GR_OK)
        return 3;
    Tally t;
    std::atomic<int> submitted{0}, refused{0};
    std::vector<std::vector<Req*>> owns(4);
    std::vector<std::thread> th;
    for (int k = 0; k < 4; ++k) {
        th.emplace_back([&, k] {
            std::mt19937 rng(100 + k);
            for (int i = 0; i < 60; ++i) {
                const int sc = (int)(rng() % 3);
                const int r = (int)(rng() % 5);
                const int tag = 1 + (int)(rng() % 250);
                if (submit(s, ids[sc], kRes[r][0], kRes[r][1], tag, t, owns[k])) ++submitted;
                else ++refused;
                if (rng() % 8 == 0) std::this_thread::sleep_for(std::chrono::microseconds(200));
            }
        });
    }
    // removal waits for the scene's in-flight requests; later submits to it are refused
    std::this_thread::sleep_for(std::chrono::milliseconds(5));
    const gr_status rm = gr_server_remove_scene(s, ids[2]);
    for (auto& x : th) x.join();
    drain(s, t, submitted.load());
    char extra[128];
    snprintf(extra, sizeof extra, "submitted=%d refused=%d remove_status=%d", submitted.load(), refused.load(), (int)rm);
    print_result(s, t, extra);
    gr_server_destroy(s);
    for (auto& o : owns)
        for (Req* r : o) delete r;
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) return 1;
    const std::string mode = argv[1];
    if (mode == "mix") return run_mix(argc > 2 ? atoi(argv[2]) : 0, argc > 3 ? atoll(argv[3]) : 0);
    if (mode == "add_budget") return run_add_budget();
    if (mode == "add_exact") return run_add_exact();
    if (mode == "add_fail") return run_add_fail();
    if (mode == "threads") return run_threads(argc > 2 ? atoi(argv[2]) : 0, argc > 3 ? atoll(argv[3]) : 0);
    return 1;
}
