// gr_overlap.hpp -- footprint rect and ellipse/tile overlap tests of the tile-list build,
// device and host (the host build checks the bounded test against the exact one).
//
// Footprint: getAABBRect core/gaussian_render.cu:274-298 and the tile rect of :379-389.
// Exact test: ellipseRectOverlap :250-272 with convertToEllipseCoordinates :234-239 and
// intersectsUnitCircle :241-248, evaluated in FP32 in the reference's operation order
// (tile_overlap).  Its result decides which tiles a Gaussian is listed in, so the build must
// reproduce it bit for bit; tile_overlap_fast returns the same result with far fewer
// instructions for all but a thin band of near-tangent tiles (overlap_bound).
#pragma once

#include "gr_math.hpp"

namespace gr {

// Footprint: getAABBRect core/gaussian_render.cu:274-298 and the tile rect of
// :379-389 (used by the projection's visibility ballot and by the build).
struct Footprint {
    int x0, y0, nx, ny;
};

GR_HD Footprint footprint(float cx, float cy, float csx, float csy, float ex,
                                               float ey, const FrameGeom& g) {
    const float rx = csx * ex, ry = csy * ex;        // principal axis
    const float ux = csy * ey, uy = (-csx) * ey;     // minor axis
    float mnx = 1e12f, mny = 1e12f, mxx = -1e12f, mxy = -1e12f;
#pragma unroll
    for (int j = 0; j != 4; ++j) {
        const int bottom = j & 1, top = j >> 1;
        const float r = (float)(bottom ^ top) * 2.0f - 1.0f;
        const float u = (float)top * 2.0f - 1.0f;
        const float vx = rx * r + ux * u;
        const float vy = ry * r + uy * u;
        mnx = gmin(mnx, vx);
        mny = gmin(mny, vy);
        mxx = gmax(mxx, vx);
        mxy = gmax(mxy, vy);
    }
    mnx = mnx + cx;
    mny = mny + cy;
    mxx = mxx + cx;
    mxy = mxy + cy;
    const float tfx = ((mnx + 1.0f) * 0.5f) * g.tiles_per_axis_x;
    const float tfy = ((mny + 1.0f) * 0.5f) * g.tiles_per_axis_y;
    const float tfz = ((mxx + 1.0f) * 0.5f) * g.tiles_per_axis_x;
    const float tfw = ((mxy + 1.0f) * 0.5f) * g.tiles_per_axis_y;
    Footprint f;
    f.x0 = iclamp(f2i_sat(floorf(tfx)), 0, g.tiles_x);
    f.y0 = iclamp(f2i_sat(floorf(tfy)), 0, g.tiles_y);
    const int x1 = iclamp(f2i_sat(ceilf(tfz)), 0, g.tiles_x);
    const int y1 = iclamp(f2i_sat(ceilf(tfw)), 0, g.tiles_y);
    f.nx = x1 - f.x0;
    f.ny = y1 - f.y0;
    return f;
}

// IEEE quotient a / b from rb = RN(1/b): q0 = RN(a*rb), e = a - b*q0 (exact
// through the FMA), q = RN(q0 + e*rb) is the correctly rounded quotient when
// nothing underflows or overflows (Markstein 1990, "Computation of elementary
// functions on the IBM RISC System/6000 processor", Theorem 1).  The caller
// guarantees 2^-40 <= |b| <= 2^40 (else fast == false); tiny or zero
// numerators take the exact paths below.  Same bits as a / b, a third of the
// instructions of the full division sequence.
GR_HD float div_by(float a, float b, float rb, bool fast) {
    if (fast && fabsf(a) >= 0x1p-40f) {
        const float q0 = a * rb;
        const float e = fmaf(-b, q0, a);
        return fmaf(e, rb, q0);
    }
    if (fast && a == 0.0f) return a * rb;
    return a / b;
}

struct EllipseRec {
    float cx, cy, csx, csy;     // centre, (cos, sin)
    float ex, ey, rex, rey;     // extents and their correctly rounded reciprocals
};

GR_HD bool div_fast_ok(float b) {
    const float ab = fabsf(b);
    return ab >= 0x1p-40f && ab <= 0x1p40f;
}

// convertToEllipseCoordinates (core/gaussian_render.cu:234-239)
GR_HD v2 to_ellipse(float px, float py, const EllipseRec& e, bool fx, bool fy) {
    const float dx = px - e.cx, dy = py - e.cy;
    return v2{div_by(dx * e.csx + dy * e.csy, e.ex, e.rex, fx), div_by(dy * e.csx - dx * e.csy, e.ey, e.rey, fy)};
}

// intersectsUnitCircle (core/gaussian_render.cu:241-248)
GR_HD bool hits_unit_circle(v2 v0, v2 v1) {
    const float ddx = v1.x - v0.x, ddy = v1.y - v0.y;
    const float lsq = ddx * ddx + ddy * ddy;
    const float t = saturate(((-v0.x) * ddx + (-v0.y) * ddy) / lsq);
    const float px = v0.x + t * (v1.x - v0.x);
    const float py = v0.y + t * (v1.y - v0.y);
    return (px * px + py * py) < 1.0f;
}

// ellipseRectOverlap (core/gaussian_render.cu:250-272) for tile (gx, gy); the
// reference ORs all six tests, here the first true one returns.
GR_HD bool tile_overlap(int gx, int gy, const EllipseRec& e, const FrameGeom& g) {
    const float mnx = (float)gx * g.tile_clip_x - 1.0f;
    const float mny = (float)gy * g.tile_clip_y - 1.0f;
    const float mxx = mnx + g.tile_clip_x;
    const float mxy = mny + g.tile_clip_y;
    if (e.cx > mnx && e.cx < mxx && e.cy > mny && e.cy < mxy) return true;
    const bool fx = div_fast_ok(e.ex), fy = div_fast_ok(e.ey);
    const float ctx = mnx + (mxx - mnx) * 0.5f;
    const float cty = mny + (mxy - mny) * 0.5f;
    if (len2sq(to_ellipse(ctx, cty, e, fx, fy)) < 1.0f) return true;
    const v2 p0 = to_ellipse(mnx, mny, e, fx, fy);
    const v2 p1 = to_ellipse(mxx, mny, e, fx, fy);
    const v2 p2 = to_ellipse(mxx, mxy, e, fx, fy);
    const v2 p3 = to_ellipse(mnx, mxy, e, fx, fy);
    return hits_unit_circle(p0, p1) || hits_unit_circle(p1, p2) || hits_unit_circle(p2, p3) ||
           hits_unit_circle(p3, p0);
}

// ---------------------------------------------------------------------------
// Bounded overlap test.
//
// In exact arithmetic the reference's test is "the ellipse and the tile rect intersect":
// the centre strictly inside the rect (its first test), or the minimum over the rect of
// Q(p) = |E(p)|^2 below 1, where E is the (linear) map convertToEllipseCoordinates applies to
// p - centre.  The minimum of the convex Q over a rect that does not contain the centre lies
// on the edges facing the centre (at most one vertical and one horizontal edge), at the
// clamped minimiser of Q along the edge's line.  overlap_bound evaluates that minimum Qmin in
// FP32 and decides when it is further than delta from 1:
//   * the reference's own FP32 evaluation (corner coordinates, closest point of each edge,
//     squared length) is within 30 eps (M + 1) of the exact answer: every point it computes
//     lies within 13 eps M of the exact segment images, and the error of its edge parameter t
//     enters only quadratically;
//   * this evaluation is within 20 eps (M + 1) of Qmin (its edge coordinates, the 1-D
//     minimiser and the map are rounded, the latter formed without cancellation);
// with eps = 2^-24 and M a bound on |E| over the record's whole footprint rect (distance from
// the centre to the farthest footprint corner, per axis, times max(1/ex, 1/ey)).  So
// Qmin < 1 - delta means the reference answers true and Qmin > 1 + delta false, for
// delta = 128 eps (M + 1).  Undecided tiles, and every tile of a record with a non-finite or
// non-positive quantity or delta >= 1/4 (very thin or tiny ellipses), go to the exact test.
// The host build of this header compares the two over random and near-tangent cases
// (tests/test_overlap_bound.py).
// ---------------------------------------------------------------------------
#ifndef GR_OVERLAP_SLACK_EPS
#define GR_OVERLAP_SLACK_EPS 128   // the host check also builds it at 0 to show it finds the band
#endif
constexpr float kOverlapSlack = (float)GR_OVERLAP_SLACK_EPS / 16777216.0f;

struct FastEllipse {
    float A, B, C, D;   // E(p) = (A dx + B dy, C dx + D dy), dx, dy = p - centre
    float kv, kh;       // minimiser of Q along x = xe: y = kv xe; along y = ye: x = kh ye
    float lo, hi;       // 1 - delta, 2 - lo = 1 + delta (-inf, +inf: always the exact test)
};

GR_HD bool finite1(float a) { return fabsf(a) <= 3.40282347e38f; }   // false for inf and NaN
GR_HD bool finite_all(float a, float b, float c, float d) {
    return finite1(a) && finite1(b) && finite1(c) && finite1(d);
}

GR_HD FastEllipse fast_ellipse(const EllipseRec& e, const Footprint& fp, float tcx, float tcy) {
    FastEllipse f;
    f.A = e.csx * e.rex;
    f.B = e.csy * e.rex;
    f.C = -e.csy * e.rey;
    f.D = e.csx * e.rey;
    const float a2 = fmaf(f.A, f.A, f.C * f.C);
    const float c2 = fmaf(f.B, f.B, f.D * f.D);
    // A B + C D = cs sn (1/ex^2 - 1/ey^2), formed from (ey - ex)(ey + ex) without cancellation
    const float b2 = ((e.csx * e.csy) * ((e.ey - e.ex) * (e.ey + e.ex))) * ((e.rex * e.rex) * (e.rey * e.rey));
    f.kv = -b2 / c2;
    f.kh = -b2 / a2;
    const float X0 = (float)fp.x0 * tcx - 1.0f, X1 = (float)(fp.x0 + fp.nx) * tcx - 1.0f;
    const float Y0 = (float)fp.y0 * tcy - 1.0f, Y1 = (float)(fp.y0 + fp.ny) * tcy - 1.0f;
    const float dm = (fmaxf(fabsf(X0 - e.cx), fabsf(X1 - e.cx)) + tcx) +
                     (fmaxf(fabsf(Y0 - e.cy), fabsf(Y1 - e.cy)) + tcy);
    const float m = dm * fmaxf(e.rex, e.rey);
    const float delta = kOverlapSlack * (m + 1.0f);
    const bool ok = e.ex > 0.0f && e.ey > 0.0f && finite_all(e.cx, e.cy, e.csx, e.csy) &&
                    finite_all(f.A, f.B, f.C, f.D) && finite_all(f.kv, f.kh, m, a2) && finite1(c2) &&
                    delta < 0.25f;
    f.lo = ok ? 1.0f - delta : -INFINITY;
    f.hi = 2.0f - f.lo;   // 1 + delta (to rounding), +inf with lo; kernels keep lo alone
    return f;
}

// The FastEllipse of a record from its EllipseRec and the (kv, kh, lo) of fast_ellipse, the
// form kernels keep per record (the map coefficients are recomputed: four products).
GR_HD FastEllipse fast_ellipse_from(const EllipseRec& e, float kv, float kh, float lo) {
    FastEllipse f;
    f.A = e.csx * e.rex;
    f.B = e.csy * e.rex;
    f.C = -e.csy * e.rey;
    f.D = e.csx * e.rey;
    f.kv = kv;
    f.kh = kh;
    f.lo = lo;
    f.hi = 2.0f - lo;
    return f;
}

// 1: the reference's test answers true, 0: false, -1: undecided (run the exact test).
GR_HD int overlap_bound(int gx, int gy, float cx, float cy, const FastEllipse& f, float tcx, float tcy) {
    // the tile rect exactly as tile_overlap forms it
    const float mnx = (float)gx * tcx - 1.0f;
    const float mny = (float)gy * tcy - 1.0f;
    const float mxx = mnx + tcx;
    const float mxy = mny + tcy;
    // signs of these differences are exact: x0 < 0 <=> cx > mnx, and so on
    const float x0 = mnx - cx, x1 = mxx - cx, y0 = mny - cy, y1 = mxy - cy;
    if (x0 < 0.0f && x1 > 0.0f && y0 < 0.0f && y1 > 0.0f) return 1;   // the reference's first test
    const bool vert = x0 > 0.0f || x1 < 0.0f;   // a vertical edge faces the centre
    const bool horz = y0 > 0.0f || y1 < 0.0f;   // a horizontal edge faces the centre
    float q = 0.0f;                              // neither: the centre is on the rect's boundary
    if (vert || horz) {
        q = INFINITY;
        if (vert) {
            const float xe = x0 > 0.0f ? x0 : x1;
            const float y = fminf(fmaxf(f.kv * xe, y0), y1);
            const float ex = fmaf(f.A, xe, f.B * y), ey = fmaf(f.C, xe, f.D * y);
            q = fmaf(ex, ex, ey * ey);
        }
        if (horz) {
            const float ye = y0 > 0.0f ? y0 : y1;
            const float x = fminf(fmaxf(f.kh * ye, x0), x1);
            const float ex = fmaf(f.A, x, f.B * ye), ey = fmaf(f.C, x, f.D * ye);
            q = fminf(q, fmaf(ex, ex, ey * ey));
        }
    }
    if (q < f.lo) return 1;
    if (q > f.hi) return 0;
    return -1;
}

// tile_overlap's result; the exact test runs only for tiles overlap_bound leaves open.
GR_HD bool tile_overlap_fast(int gx, int gy, const EllipseRec& e, const FastEllipse& f, const FrameGeom& g) {
    const int d = overlap_bound(gx, gy, e.cx, e.cy, f, g.tile_clip_x, g.tile_clip_y);
    return d < 0 ? tile_overlap(gx, gy, e, g) : d != 0;
}

}  // namespace gr
