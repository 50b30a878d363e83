// overlap_check.cpp -- the bounded overlap test (tile_overlap_fast, gr_overlap.hpp) against the
// reference's exact FP32 test (tile_overlap), host build of the same header.  Test
// infrastructure only.
//
//   overlap_check <mode> <seed> <records>
//     mode "random":  ellipses as the projection makes them (any orientation, extents from a
//                     hundredth of a pixel to the whole screen, aspect ratios to 1e-4, circles),
//                     every candidate tile of each footprint (up to 256 per record)
//     mode "tangent": each ellipse rescaled so that one chosen tile is within 1e-4 .. 1e-8 of
//                     tangency (relative), then that tile and its neighbours
//     mode "special": centres on tile edges and corners, zero/denormal/huge/non-finite extents,
//                     zero and non-unit direction vectors
// Prints "RESULT tests=<n> mismatch=<n> undecided=<n>".
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>

#include "gr_overlap.hpp"

using namespace gr;

namespace {

const int kSizes[6][2] = {{1920, 1080}, {3840, 2160}, {800, 600}, {640, 360}, {333, 257}, {1024, 1024}};

FrameGeom geom_for(int w, int h) { return make_geom(w, h); }

struct Tally {
    long long tests = 0, mismatch = 0, undecided = 0;
    void check(int gx, int gy, const EllipseRec& e, const FastEllipse& f, const FrameGeom& g) {
        const bool exact = tile_overlap(gx, gy, e, g);
        const int d = overlap_bound(gx, gy, e.cx, e.cy, f, g.tile_clip_x, g.tile_clip_y);
        const bool fast = tile_overlap_fast(gx, gy, e, f, g);
        ++tests;
        if (d < 0) ++undecided;
        if (fast != exact) {
            if (mismatch < 10)
                std::printf("MISMATCH tile (%d,%d) exact %d fast %d bound %d: c=(%.9g,%.9g) cs=(%.9g,%.9g) "
                            "ext=(%.9g,%.9g) lo=%.9g hi=%.9g\n",
                            gx, gy, (int)exact, (int)fast, d, e.cx, e.cy, e.csx, e.csy, e.ex, e.ey, f.lo, f.hi);
            ++mismatch;
        }
    }
};

EllipseRec make_rec(float cx, float cy, float csx, float csy, float ex, float ey) {
    return EllipseRec{cx, cy, csx, csy, ex, ey, 1.0f / ex, 1.0f / ey};
}

// Exact-math minimum over tile (gx, gy) of |E|^2, in double (to place near-tangent cases).
double qmin_double(int gx, int gy, const EllipseRec& e, const FrameGeom& g) {
    const float mnx = (float)gx * g.tile_clip_x - 1.0f, mny = (float)gy * g.tile_clip_y - 1.0f;
    const double x0 = (double)mnx - e.cx, x1 = (double)(mnx + g.tile_clip_x) - e.cx;
    const double y0 = (double)mny - e.cy, y1 = (double)(mny + g.tile_clip_y) - e.cy;
    const double A = e.csx / (double)e.ex, B = e.csy / (double)e.ex, C = -e.csy / (double)e.ey,
                 D = e.csx / (double)e.ey;
    auto Q = [&](double x, double y) {
        const double u = A * x + B * y, v = C * x + D * y;
        return u * u + v * v;
    };
    if (x0 <= 0 && x1 >= 0 && y0 <= 0 && y1 >= 0) return 0.0;
    double best = 1e300;
    // dense search refined along the boundary (the minimum lies on the boundary)
    for (int side = 0; side < 4; ++side) {
        double lo = 0.0, hi = 1.0;
        auto pt = [&](double t, double& x, double& y) {
            if (side == 0) { x = x0 + t * (x1 - x0); y = y0; }
            else if (side == 1) { x = x1; y = y0 + t * (y1 - y0); }
            else if (side == 2) { x = x0 + t * (x1 - x0); y = y1; }
            else { x = x0; y = y0 + t * (y1 - y0); }
        };
        for (int it = 0; it < 200; ++it) {   // ternary search on a convex function
            const double m1 = lo + (hi - lo) / 3, m2 = hi - (hi - lo) / 3;
            double xa, ya, xb, yb;
            pt(m1, xa, ya);
            pt(m2, xb, yb);
            if (Q(xa, ya) < Q(xb, yb)) hi = m2; else lo = m1;
        }
        double x, y;
        pt(0.5 * (lo + hi), x, y);
        best = std::fmin(best, Q(x, y));
    }
    return best;
}

void check_record(Tally& t, const EllipseRec& e, const FrameGeom& g, int max_tiles) {
    const Footprint fp = footprint(e.cx, e.cy, e.csx, e.csy, e.ex, e.ey, g);
    const FastEllipse f = fast_ellipse(e, fp, g.tile_clip_x, g.tile_clip_y);
    if (fp.nx <= 0 || fp.ny <= 0) return;
    const long long cand = (long long)fp.nx * fp.ny;
    const long long step = cand > max_tiles ? cand / max_tiles + 1 : 1;
    for (long long c = 0; c < cand; c += step)
        t.check(fp.x0 + (int)(c % fp.nx), fp.y0 + (int)(c / fp.nx), e, f, g);
}

float log_uniform(std::mt19937_64& r, double lo, double hi) {
    std::uniform_real_distribution<double> u(std::log(lo), std::log(hi));
    return (float)std::exp(u(r));
}

void unit_dir(std::mt19937_64& r, float& cs, float& sn) {
    std::uniform_real_distribution<double> a(-3.14159265358979, 3.14159265358979);
    const double th = a(r);
    const v2 e0 = normalize2(v2{(float)std::cos(th), (float)std::sin(th)});
    cs = e0.x;
    sn = e0.y;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 4) return 1;
    const std::string mode = argv[1];
    std::mt19937_64 rng((uint64_t)std::atoll(argv[2]));
    const long long records = std::atoll(argv[3]);
    std::uniform_real_distribution<double> u01(0.0, 1.0);
    Tally t;
    for (long long i = 0; i < records; ++i) {
        const int* wh = kSizes[rng() % 6];
        const FrameGeom g = geom_for(wh[0], wh[1]);
        const float px = 2.0f / (float)wh[0];
        float cs, sn;
        unit_dir(rng, cs, sn);
        const float cx = (float)(u01(rng) * 2.0 - 1.0), cy = (float)(u01(rng) * 2.0 - 1.0);
        float ex = log_uniform(rng, 0.01 * px, 2.0);
        float ey = ex * ((rng() % 8 == 0) ? 1.0f : log_uniform(rng, 1e-4, 1.0));
        if (mode == "random") {
            check_record(t, make_rec(cx, cy, cs, sn, ex, ey), g, 256);
        } else if (mode == "tangent") {
            const EllipseRec e0 = make_rec(cx, cy, cs, sn, ex, ey);
            const Footprint fp = footprint(e0.cx, e0.cy, e0.csx, e0.csy, e0.ex, e0.ey, g);
            if (fp.nx <= 0 || fp.ny <= 0) continue;
            const int gx = fp.x0 + (int)(rng() % fp.nx), gy = fp.y0 + (int)(rng() % fp.ny);
            const double q = qmin_double(gx, gy, e0, g);
            if (!(q > 0.0)) continue;
            // scale both extents so that this tile sits at Qmin ~ 1 - 2u
            const double uu = (u01(rng) * 2.0 - 1.0) * std::pow(10.0, -4.0 - 4.0 * u01(rng));
            const double s = std::sqrt(q) * (1.0 + uu);
            const EllipseRec e = make_rec(cx, cy, cs, sn, (float)(ex * s), (float)(ey * s));
            const Footprint f2 = footprint(e.cx, e.cy, e.csx, e.csy, e.ex, e.ey, g);
            const FastEllipse f = fast_ellipse(e, f2, g.tile_clip_x, g.tile_clip_y);
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) t.check(gx + dx, gy + dy, e, f, g);
        } else if (mode == "special") {
            const int k = (int)(rng() % 8);
            float scx = cx, scy = cy;
            if (k == 0 || k == 1) {   // centre exactly on a tile edge / corner
                const int gx = (int)(rng() % g.tiles_x), gy = (int)(rng() % g.tiles_y);
                scx = (float)gx * g.tile_clip_x - 1.0f;
                if (k == 1) scy = (float)gy * g.tile_clip_y - 1.0f;
            }
            if (k == 2) ey = 0.0f;
            if (k == 3) ex = 1e-38f;
            if (k == 4) ey = 1e30f;
            if (k == 5) { cs = 0.0f; sn = 0.0f; }
            if (k == 6) { cs *= 1.5f; sn *= 0.5f; }
            if (k == 7) ex = (rng() % 2) ? NAN : INFINITY;
            check_record(t, make_rec(scx, scy, cs, sn, ex, ey), g, 64);
        } else {
            return 1;
        }
    }
    std::printf("RESULT tests=%lld mismatch=%lld undecided=%lld\n", t.tests, t.mismatch, t.undecided);
    return 0;
}
