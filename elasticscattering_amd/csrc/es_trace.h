// es_trace.h — device side of one particle's lifetime (gfx950, FP64).
//
// The floating-point operations and their order are those of the reference
// kernel library (src/sim/es/details.h, cell_grid.h, lifetime.h); the code is
// organised for the GPU:
//   * per-orbit constants are computed once per particle;
//   * the rotation sense is a template parameter;
//   * the four cell edges and the two crossing points are loops (one copy of
//     the asin-bearing code, which keeps the register count down);
//   * a trace advances one cell per call of d_step(), so a wave can give a
//     lane a new particle as soon as its trace ends (es_kernels.hip);
//   * impurities are read as 16-byte (x, y) pairs, four cells' worth of
//     loads in flight;
//   * the start-cell "inside an impurity" test is hoisted out: it depends on
//     neither phi nor B and runs once per position (k_positions).
// Compiled with -ffp-contract=off so no multiply-add is fused.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "es_math.h"

// ES_TRACE_SCREEN=1 solves the crossing disks of a cell in order of a lower
// bound on their travel angle and skips those that cannot lower the minimum
// (results identical).  Per-lane skips only pay when a whole wave skips, so it
// is off by default.
#ifndef ES_TRACE_SCREEN
#define ES_TRACE_SCREEN 0
#endif

#define ES_PI 3.141592653589793238463
#define ES_PI2 6.283185307179586
#define ES_HALF_PI 1.570796326794896619231
#define ES_INF 1.7976931348623157e308   /* DBL_MAX (constants.h:36) */
#define ES_INV_PI2 (1.0 / ES_PI2)          /* RN(1 / 2 pi), folded by the compiler */
#define ES_INV_HALF_PI (1.0 / ES_HALF_PI)  /* RN(2 / pi) */

// Device view of the binned impurity field (Grid::GetSettings, Grid.h:82-91).
struct DevGrid {
    const double2* ixy;       // cell-ordered (x, y) of indexed impurities
    const int32_t* cell_end;  // inclusive end offset per cell (Grid::imp_index)
    double start;             // spawn_region_start
    double size;              // spawn_region_size
    double L;                 // cell_size
    double inv_L;             // RN(1 / L), for esm_div_rcp
    double radius;            // impurity_radius
    int cpr;                  // cells_per_row
};

// One orbit (details.h:31-65) with the constants its tests reuse.
struct Orb {
    double cx, cy;          // centre
    double R, R2;           // radius, radius^2
    double eps_r;           // 1e-10 * R (details.h:98)
    double radd2, rmin2;    // (R + r)^2, (R - r)^2 (details.h:147-150)
    double r2;              // r * r (details.h:160)
    double pa;              // particle_angle = phi
    double w;               // angular speed (unsigned)
    double inv_R, inv_w;    // RN(1 / R), RN(1 / w), for esm_div_rcp
    double u0x, u0y;        // start position minus centre (screening only)
    float r;                // impurity radius (screening only)
};

// Cell-boundary crossing (cell_grid.h:31-46); dphi lives in the search only.
struct Isect {
    double x, y;
    double ang;             // incident angle
    int cx, cy;             // entering cell
};

__device__ __forceinline__ double d_smod(double a, double b)
{
    return a - b * floor(a / b);  // details.h:79-83
}

// smod(a, 2 pi) with the quotient from the precomputed reciprocal (identical bits).
__device__ __forceinline__ double d_smod_2pi(double a)
{
    return a - ES_PI2 * floor(esm_div_rcp(a, ES_PI2, ES_INV_PI2));
}

// GetAngle + AngleVelocity (details.h:85-107).
template <bool CW>
__device__ __forceinline__ double d_angle(double x, double y, const Orb& o)
{
    const double y_dist = y - o.cy;
    const bool left_side = x < o.cx;
    double angle;
    if (fabs(y_dist) > o.eps_r) {
        angle = esm_asin(esm_div_rcp(y_dist, o.R, o.inv_R));
        angle = left_side ? ES_PI - angle : angle;
    } else {
        angle = left_side ? ES_PI : 0.0;
    }
    return d_smod_2pi(angle + (CW ? ES_HALF_PI : -ES_HALF_PI));
}

// d_angle for the paths that rarely run (undecided pseudo-angle margins).
// Inlined: an out-of-line copy measured 1.7 % slower on the uneven batch
// (call overhead and spills around the call) and no faster on the dense field.
template <bool CW>
__device__ __forceinline__ double d_angle_rare(double x, double y, double cx, double cy, double R, double inv_R,
                                               double eps_r)
{
    Orb o;
    o.cx = cx;
    o.cy = cy;
    o.R = R;
    o.inv_R = inv_R;
    o.eps_r = eps_r;
    return d_angle<CW>(x, y, o);
}

// GetCrossAngle (details.h:111-116).
template <bool CW>
__device__ __forceinline__ double d_cross_angle(double p, double q)
{
    return d_smod_2pi(CW ? (p - q) : (q - p));
}

// AngleInRange (details.h:119-136).
template <bool CW>
__device__ __forceinline__ bool d_in_range(double phi, double rx, double ry)
{
    if (fabs(rx - ry) < 1e-10) return true;
    const double low = CW ? ry : rx;
    double high = CW ? rx : ry;
    high = (high < low) ? (high + ES_PI2) : high;
    return (low < phi && phi < high) || (low < (phi + ES_PI2) && (phi + ES_PI2) < high);
}

// ------------------------------------------------------ travel pseudo-angles
//
// Not reference operations: a cheap order on points of the orbit, used to
// decide which points need their angle (asin) at all.  The travel angle from
// a point A to a point X of the orbit (both as vectors from the centre) is
// the angle swept in the direction of motion, in [0, 2 pi).  Its
// pseudo-angle, from the dot and cross products of the two vectors, is in
// [0, 4), increases with it, and changes by at least the change of the
// travel angle times 1/2 and at most the change itself (slope in [1/2, 1]).
// So two pseudo-angles more than m apart belong to travel angles more than m
// apart.
//
// The reference's angle of a point (GetAngle, details.h:93-107) is asin of
// (y - cy) / R.  For a point within a few ulps of the circle (every crossing
// and edge point the reference forms is) and with |y - cy| < (1 - 1e-10) R,
// asin's slope is below 7.1e4, so that angle is the point's true angle to
// about 3e-11 rad, and to 1e-10 rad in the |y - cy| <= 1e-10 R branch.  Every
// comparison below keeps a margin of at least 1e-9 (1e-6 against the FP32
// cell-range values), so an order decided here is the order the reference's
// computed angles have; where a margin or the |y - cy| bound fails, nothing
// is decided and the reference operations run.

constexpr double kTravelMargin = 1e-9;             // between two FP64 pseudo-angles
constexpr double kRangeMargin = 1e-6;              // against a cell range bound stored as FP32
constexpr double kAngleSlopeLimit = 0.9999999999;  // |y - cy| / R below this: asin's slope < 7.1e4
constexpr double kTurnMin = 1e-6;                  // a turning point nearer the start is taken by bounds only

__device__ __forceinline__ double d_pseudo_angle(double x, double y)
{
    const double ax = fabs(x), ay = fabs(y);
    const double sum = ax + ay;
    // ay / sum by v_rcp_f64 and one Newton step (relative error below 1e-14;
    // an IEEE division is three times the instructions).  Every use compares
    // pseudo-angles on margins of 1e-9 or more.  sum >= |A| |V| is a normal
    // number for any orbit the solver accepts.
    double q = __builtin_amdgcn_rcp(sum);
    q = fma(fma(-sum, q, 1.0), q, q);
    const double r = (sum > 0.0) ? ay * q : 0.0;
    const bool up = y >= 0.0;
    const bool plus = up ? (x >= 0.0) : (x < 0.0);
    const double base = up ? (plus ? 0.0 : 2.0) : (plus ? 2.0 : 4.0);
    return plus ? base + r : base - r;
}

// Pseudo-angle of the travel angle from A = (ax, ay) to V = (vx, vy) (vectors
// from the orbit centre).
template <bool CW>
__device__ __forceinline__ double d_travel_pseudo(double ax, double ay, double vx, double vy)
{
    const double dot = ax * vx + ay * vy;
    const double crs = ax * vy - ay * vx;
    return d_pseudo_angle(dot, CW ? -crs : crs);
}

// GetFirstCrossTime with GetCrossPoints inlined (details.h:153-206).
template <bool CW>
__device__ __forceinline__ double d_first_cross_time(const Orb& o, double px, double py, double rx, double ry)
{
    const double qx = o.cx - px, qy = o.cy - py;
    const double dist_squared = qx * qx + qy * qy;
    const double dist = sqrt(dist_squared);
    const double inv_d = 1.0 / dist;  // RN(1 / (2 dist)) = RN(1 / dist) / 2
    const double xs = esm_div_rcp(dist_squared + o.R2 - o.r2, 2.0 * dist, 0.5 * inv_d);
    const double ys = sqrt(o.R2 - xs * xs);
    const double ux = esm_div_rcp(px - o.cx, dist, inv_d), uy = esm_div_rcp(py - o.cy, dist, inv_d);
    const double bx = o.cx + ux * xs, by = o.cy + uy * xs;
    double tt = ES_INF;
#pragma unroll 1
    for (int k = 0; k < 2; k++) {
        // point 1: (bx + uy*ys, by + (-ux)*ys); point 2: (bx + (-uy)*ys, by + ux*ys)
        const double ex = k ? -uy : uy;
        const double ey = k ? ux : -ux;
        const double phi_k = d_angle<CW>(bx + ex * ys, by + ey * ys, o);
        if (d_in_range<CW>(phi_k, rx, ry)) {
            const double t = esm_div_rcp(d_cross_angle<CW>(o.pa, phi_k), o.w, o.inv_w);
            tt = (tt < t) ? tt : t;
        }
    }
    return tt;
}

// d_first_cross_time with the angles taken only where they can matter (same
// result).  (u0x, u0y): the start point minus the centre; pe, pf: the cell's
// accepted range as travel pseudo-angles from the start (cell_range_pseudo:
// pe = -1, pf = 5 when AngleInRange accepts every angle, pe NaN when not
// known well enough).  A crossing point is in the range when its travel
// pseudo-angle p lies in (pe, pf) by the margin, out of it when outside by
// the margin, undecided otherwise or when p is within the margin of a whole
// turn (its angle could round to either side of the start).  Of two points
// in range only the earlier can give the minimum time (the time is the
// travel angle over w).  Decided points out of range and the later point are
// not evaluated, and a point decided in range takes its time without
// AngleInRange; only when something is undecided do both points run the
// reference operations, with the cell's range (rx, ry) as the reference forms
// it: the angles of the entry (ex, ey) and exit (fx, fy) points, or the start
// angle o.pa for a NaN ex (the first cell) or fx (no next cell).
template <bool CW>
__device__ __forceinline__ double d_first_cross_time_sel(const Orb& o, double px, double py, double ex, double ey,
                                                         double fx, double fy, double u0x, double u0y, float pe,
                                                         float pf)
{
    const double qx = o.cx - px, qy = o.cy - py;
    const double dist_squared = qx * qx + qy * qy;
    const double dist = sqrt(dist_squared);
    const double inv_d = 1.0 / dist;  // RN(1 / (2 dist)) = RN(1 / dist) / 2
    const double xs = esm_div_rcp(dist_squared + o.R2 - o.r2, 2.0 * dist, 0.5 * inv_d);
    const double ys = sqrt(o.R2 - xs * xs);
    const double ux = esm_div_rcp(px - o.cx, dist, inv_d), uy = esm_div_rcp(py - o.cy, dist, inv_d);
    const double bx = o.cx + ux * xs, by = o.cy + uy * xs;
    unsigned int need = 3u;  // bit k: point k runs the reference operations
    bool decided = false;    // every point of need decided in range
    {
        const double lim = kAngleSlopeLimit * o.R;
        bool ok = (ys == ys) && (pe == pe);
        double p[2];
        unsigned int in = 0u, out = 0u;
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const double ex = k ? -uy : uy;
            const double ey = k ? ux : -ux;
            const double vx = (bx + ex * ys) - o.cx, vy = (by + ey * ys) - o.cy;
            ok = ok && fabs(vy) < lim;
            p[k] = d_travel_pseudo<CW>(u0x, u0y, vx, vy);
            const bool turn_ok = p[k] > kTravelMargin && p[k] < 4.0 - kTravelMargin;
            in |= (turn_ok && p[k] > (double)pe + kRangeMargin && p[k] < (double)pf - kRangeMargin) ? (1u << k) : 0u;
            out |= (turn_ok && (p[k] < (double)pe - kRangeMargin || p[k] > (double)pf + kRangeMargin)) ? (1u << k) : 0u;
        }
        if (ok && (in | out) == 3u) {
            need = in;
            decided = true;
            if (in == 3u) {
                need = (p[0] + kTravelMargin < p[1]) ? 1u : (p[1] + kTravelMargin < p[0]) ? 2u : 3u;
                decided = need != 3u;
            }
        }
    }
    double tt = ES_INF;
    double rx = 0.0, ry = 0.0;
    if (!decided) {  // the cell's range as the reference forms it (lifetime.h:106-115)
        rx = (ex == ex) ? d_angle_rare<CW>(ex, ey, o.cx, o.cy, o.R, o.inv_R, o.eps_r) : o.pa;
        ry = (fx == fx) ? d_angle_rare<CW>(fx, fy, o.cx, o.cy, o.R, o.inv_R, o.eps_r) : o.pa;
    }
    while (need) {  // one copy of the angle code
        const int k = __builtin_ctz(need);
        need &= need - 1u;
        const double dx = k ? -uy : uy;
        const double dy = k ? ux : -ux;
        const double phi_k = d_angle<CW>(bx + dx * ys, by + dy * ys, o);
        if (decided || d_in_range<CW>(phi_k, rx, ry)) {
            const double t = esm_div_rcp(d_cross_angle<CW>(o.pa, phi_k), o.w, o.inv_w);
            tt = (tt < t) ? tt : t;
        }
    }
    return tt;
}

// The cell's edges as UpdateFirstBoundaryIntersect (cell_grid.h:91-129) sees
// them.  With low_left = (X0, Y0) the corners are (X0, Y0), (X1, Y0), (X1, Y1),
// (X0, Y1), X1 = X0 + L, Y1 = Y0 + L (cell_grid.h:140-143); the edge direction
// u = (p2 - p1) / L is (qx, 0) for the horizontal edges and (0, qy) for the
// vertical ones (0 / L = +0), qx = (X1 - X0) / L and qy = (Y1 - Y0) / L.  Both
// horizontal edges start at x = X0, so their projection distance
// u . (centre - p1) = qx (cx - X0) + 0 (cy - Y) and projected x are one value,
// and both vertical edges share the projected y: adding the +-0 of the zero
// component changes at most the sign of a zero, which no later comparison or
// difference of the path sees.  PointInSegment(a, l0, l1) (cell_grid.h:69-72)
// is l0 < a < l1 here, as l0 < l1.
struct CellEdges {
    double X0, X1, Y0, Y1;  // edge lines
    double qx, qy;          // edge directions
    double projx, projy;    // projection of the centre onto the horizontal / vertical edge lines
    double dxh, dyv;        // projx - cx, projy - cy
};

__device__ __forceinline__ void d_cell_edges(const Orb& o, const DevGrid& g, int cx, int cy, CellEdges& c)
{
    const double L = g.L;
    c.X0 = g.start + (double)cx * L;  // to_world (cell_grid.h:60-63)
    c.Y0 = g.start + (double)cy * L;
    c.X1 = c.X0 + L;
    c.Y1 = c.Y0 + L;
    c.qx = esm_div_rcp(c.X1 - c.X0, L, g.inv_L);
    c.qy = esm_div_rcp(c.Y1 - c.Y0, L, g.inv_L);
    c.projx = c.X0 + c.qx * (c.qx * (o.cx - c.X0));
    c.projy = c.Y0 + c.qy * (c.qy * (o.cy - c.Y0));
    c.dxh = c.projx - o.cx;
    c.dyv = c.projy - o.cy;
}

// Edge e in the reference order — top (0,+1), bottom (0,-1), right (+1,0),
// left (-1,0): false when its line misses the circle, else the half chord t
// (|u| sqrt(R^2 - distance^2), signed by u) and the projected coordinate along
// the edge; point k lies at base + t (k = 0) or base - t (k = 1).
__device__ __forceinline__ bool d_edge_chord(const Orb& o, const CellEdges& c, int e, double& base, double& t)
{
    const bool h = e < 2;
    const double dx = h ? c.dxh : (((e == 2) ? c.X1 : c.X0) - o.cx);
    const double dy = h ? (((e == 0) ? c.Y1 : c.Y0) - o.cy) : c.dyv;
    const double pcd = dx * dx + dy * dy;
    if (pcd >= o.R2) return false;
    t = (h ? c.qx : c.qy) * sqrt(o.R2 - pcd);
    base = h ? c.projx : c.projy;
    return true;
}

// Point k of edge e from its chord: (a, Y) on a horizontal edge, (X, a) on a vertical one.
__device__ __forceinline__ void d_edge_xy(const CellEdges& c, int e, int k, double base, double t, double& ix,
                                          double& iy)
{
    const double a = k ? (base - t) : (base + t);
    const bool h = e < 2;
    ix = h ? a : ((e == 2) ? c.X1 : c.X0);
    iy = h ? ((e == 0) ? c.Y1 : c.Y0) : a;
}

// GetNextCell (cell_grid.h:132-181) with UpdateFirstBoundaryIntersect
// (:91-129) and Update_Best_Intersect (:80-89).  `phi0`: the start angle.
// A first pass over the four edges finds which of the 8 edge-line points lie
// in their segment and differ from the entry point (DifferentPoint, :74-77: a
// point it rejects is rejected whatever its angle); a second pass takes only
// those, in the reference's order, for their angles (asin) and the smallest
// dphi — so a wave runs the angle once or twice per cell step, not once per
// point any lane has.
template <bool CW>
__device__ __forceinline__ bool d_next_cell(const Orb& o, double phi0, const DevGrid& g, const Isect& last, Isect& next,
                                            int& edge_hits)
{
    double best_dphi = ES_PI2;
    next = last;
    const double L = g.L;
    const double dp_eps = 1e-6 * L;
    CellEdges c;
    d_cell_edges(o, g, last.cx, last.cy, c);
    const double llx = c.X0, lly = c.Y0;

    unsigned int cand = 0;  // bit 2 e + k: a point in its segment, not the entry point
#pragma unroll
    for (int e = 0; e < 4; e++) {
        double base, t;
        if (!d_edge_chord(o, c, e, base, t)) continue;
        const double l0 = (e < 2) ? c.X0 : c.Y0, l1 = (e < 2) ? c.X1 : c.Y1;
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const double a = k ? (base - t) : (base + t);
            if (!(a > l0 && a < l1)) continue;
            edge_hits += 1;
            double ix, iy;
            d_edge_xy(c, e, k, base, t, ix, iy);
            if (fabs(ix - last.x) > dp_eps || fabs(iy - last.y) > dp_eps) cand |= 1u << (2 * e + k);
        }
    }
    while (cand) {
        const int bit = __builtin_ctz(cand);
        cand &= cand - 1u;
        const int e = bit >> 1, k = bit & 1;
        double base = 0.0, t = 0.0, ix, iy;
        (void)d_edge_chord(o, c, e, base, t);  // the same operations: the same point
        d_edge_xy(c, e, k, base, t, ix, iy);
        const double ia = d_angle<CW>(ix, iy, o);
        const double dphi = d_cross_angle<CW>(last.ang, ia);
        if (dphi < best_dphi) {
            next.x = ix;
            next.y = iy;
            next.cx = last.cx + ((e == 2) ? 1 : (e == 3) ? -1 : 0);
            next.cy = last.cy + ((e == 0) ? 1 : (e == 1) ? -1 : 0);
            next.ang = ia;
            best_dphi = dphi;
        }
    }

    const double delta = 1e-7 * L;
    const bool is_xlow = (next.x - llx) < delta;
    const bool is_xtop = (-next.x + L + llx) < delta;
    const bool is_ylow = (next.y - lly) < delta;
    const bool is_ytop = (-next.y + L + lly) < delta;
    int ox = next.cx - last.cx, oy = next.cy - last.cy;
    if (is_xtop && is_ylow) { ox = 1; oy = -1; }
    if (is_xlow && is_ylow) { ox = -1; oy = -1; }
    if (is_xlow && is_ytop) { ox = -1; oy = 1; }
    if (is_xtop && is_ytop) { ox = 1; oy = 1; }
    next.cx = last.cx + ox;
    next.cy = last.cy + oy;

    const bool same_cell = next.cx == last.cx && next.cy == last.cy;
    const bool inside = next.cx >= 0 && next.cx < g.cpr && next.cy >= 0 && next.cy < g.cpr;
    return !same_cell && inside && !d_in_range<CW>(phi0, last.ang, next.ang);
}

// d_next_cell without the angles where they cannot matter (same exit, same
// next cell, same availability).  `last.ang` may be NaN: the entry's angle not
// taken yet (it is only needed below when something is undecided); `first`:
// the first cell, whose entry is the start point and whose entry angle is
// phi0 itself.  The exit is the candidate with the smallest travel pseudo-angle
// from the entry (GetNextCell keeps the smallest dphi, cell_grid.h:80-89),
// decided when every candidate and the entry lie within the slope bound, the
// winner is more than the margin from the entry and from a whole turn, and
// the runner-up more than the margin behind it; its angle is then left NaN
// in `next.ang`.  "The next arc holds the start angle" (cell_grid.h:177-180)
// is the exit's travel from the start lying before the entry's (a later cell)
// — false in a first cell except counter-clockwise from a raw start angle
// phi0 < 0, where it holds while the exit comes before the top of the orbit
// (travel -phi0, where the velocity angle is 0) — decided by margins the same
// way.  Anything undecided takes the reference operations, with the angles
// they need formed from the same points (the same bits).
template <bool CW>
__device__ __forceinline__ bool d_next_cell_lazy(const Orb& o, double phi0, bool first, const DevGrid& g,
                                                 const Isect& last, Isect& next, int& edge_hits, double& pe_start,
                                                 double& pf_start)
{
    pe_start = pf_start = __builtin_nan("");
    next = last;
    next.ang = __builtin_nan("");
    const double L = g.L;
    const double dp_eps = 1e-6 * L;
    CellEdges c;
    d_cell_edges(o, g, last.cx, last.cy, c);
    const double llx = c.X0, lly = c.Y0;
    const double lim = kAngleSlopeLimit * o.R;

    unsigned int cand = 0;  // bit 2 e + k: a point in its segment, not the entry point
#pragma unroll
    for (int e = 0; e < 4; e++) {
        double base, t;
        if (!d_edge_chord(o, c, e, base, t)) continue;
        const double l0 = (e < 2) ? c.X0 : c.Y0, l1 = (e < 2) ? c.X1 : c.Y1;
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const double a = k ? (base - t) : (base + t);
            if (!(a > l0 && a < l1)) continue;
            edge_hits += 1;
            double ix, iy;
            d_edge_xy(c, e, k, base, t, ix, iy);
            if (fabs(ix - last.x) > dp_eps || fabs(iy - last.y) > dp_eps) cand |= 1u << (2 * e + k);
        }
    }
    // The candidates by travel pseudo-angle from the entry: the smallest and
    // the runner-up.
    const unsigned int cand0 = cand;
    bool ok = first || fabs(last.y - o.cy) < lim;
    double pb = 5.0, ps = 5.0;
    while (cand) {
        const int bit = __builtin_ctz(cand);
        cand &= cand - 1u;
        const int e = bit >> 1, k = bit & 1;
        double base = 0.0, t = 0.0, ix, iy;
        (void)d_edge_chord(o, c, e, base, t);  // the same operations: the same point
        d_edge_xy(c, e, k, base, t, ix, iy);
        const double vy = iy - o.cy;
        ok = ok && fabs(vy) < lim;
        const double p = d_travel_pseudo<CW>(last.x - o.cx, last.y - o.cy, ix - o.cx, vy);
        if (p < pb) {
            ps = pb;
            pb = p;
            next.x = ix;
            next.y = iy;
            next.cx = last.cx + ((e == 2) ? 1 : (e == 3) ? -1 : 0);
            next.cy = last.cy + ((e == 0) ? 1 : (e == 1) ? -1 : 0);
        } else if (p < ps) {
            ps = p;
        }
    }
    const bool decided = ok && (cand0 == 0u || (pb > kTravelMargin && pb < 4.0 - kTravelMargin && ps - pb > kTravelMargin));
    double last_ang = last.ang;
    if (!decided) {  // the reference's choice: the smallest dphi from the entry's angle
        next = last;
        if (last_ang != last_ang) last_ang = d_angle_rare<CW>(last.x, last.y, o.cx, o.cy, o.R, o.inv_R, o.eps_r);
        double best_dphi = ES_PI2;
        cand = cand0;
        while (cand) {
            const int bit = __builtin_ctz(cand);
            cand &= cand - 1u;
            const int e = bit >> 1, k = bit & 1;
            double base = 0.0, t = 0.0, ix, iy;
            (void)d_edge_chord(o, c, e, base, t);
            d_edge_xy(c, e, k, base, t, ix, iy);
            const double ia = d_angle_rare<CW>(ix, iy, o.cx, o.cy, o.R, o.inv_R, o.eps_r);
            const double dphi = d_cross_angle<CW>(last_ang, ia);
            if (dphi < best_dphi) {
                next.x = ix;
                next.y = iy;
                next.cx = last.cx + ((e == 2) ? 1 : (e == 3) ? -1 : 0);
                next.cy = last.cy + ((e == 0) ? 1 : (e == 1) ? -1 : 0);
                next.ang = ia;
                best_dphi = dphi;
            }
        }
    }

    const double delta = 1e-7 * L;
    const bool is_xlow = (next.x - llx) < delta;
    const bool is_xtop = (-next.x + L + llx) < delta;
    const bool is_ylow = (next.y - lly) < delta;
    const bool is_ytop = (-next.y + L + lly) < delta;
    int ox = next.cx - last.cx, oy = next.cy - last.cy;
    if (is_xtop && is_ylow) { ox = 1; oy = -1; }
    if (is_xlow && is_ylow) { ox = -1; oy = -1; }
    if (is_xlow && is_ytop) { ox = -1; oy = 1; }
    if (is_xtop && is_ytop) { ox = 1; oy = 1; }
    next.cx = last.cx + ox;
    next.cy = last.cy + oy;

    const bool same_cell = next.cx == last.cx && next.cy == last.cy;
    const bool inside = next.cx >= 0 && next.cx < g.cpr && next.cy >= 0 && next.cy < g.cpr;
    if (same_cell || !inside) return false;
    if (decided) {
        const double fvy = next.y - o.cy;
        const double f = d_travel_pseudo<CW>(o.u0x, o.u0y, next.x - o.cx, fvy);
        pf_start = f;
        bool dec = fabs(fvy) < lim && f > kRangeMargin && f < 4.0 - kRangeMargin;
        bool held = false;
        if (first) {
            if (!CW && phi0 < 0.0) {  // the top of the orbit, travel m = -phi0: measured only from kTurnMin on
                const double pt = d_travel_pseudo<CW>(o.u0x, o.u0y, 0.0, o.R);
                dec = dec && -phi0 >= kTurnMin && fabs(f - pt) > kRangeMargin;
                held = f < pt;
            }
        } else {
            const double evy = last.y - o.cy;
            const double e = d_travel_pseudo<CW>(o.u0x, o.u0y, last.x - o.cx, evy);
            pe_start = e;
            dec = dec && fabs(evy) < lim && e > kRangeMargin && e < 4.0 - kRangeMargin && fabs(f - e) > kRangeMargin;
            held = f < e;
        }
        if (dec) return !held;
        if (last_ang != last_ang) last_ang = d_angle_rare<CW>(last.x, last.y, o.cx, o.cy, o.R, o.inv_R, o.eps_r);
        next.ang = d_angle_rare<CW>(next.x, next.y, o.cx, o.cy, o.R, o.inv_R, o.eps_r);
    }
    return !d_in_range<CW>(phi0, last_ang, next.ang);
}

// Screening bound (not a reference operation): a lower bound on the travel
// angle, in the direction of motion, from the start point P0 to any point where
// the orbit meets the disk of radius r around C = (px, py), or 0 when no bound
// is available.
//  * The crossing points lie within asin(r/d) <= 1.0473 r/d (r/d <= 1/2) of
//    the direction of C seen from the orbit centre O (tangent lines from O).
//  * If |P0 - C| >= r + 1e-6 R, P0 is outside the disk, every crossing point X
//    has |P0 - X| >= |P0 - C| - r >= 1e-6 R, so its travel angle lies in
//    [1e-6, 2pi - 1e-6]: the FP64 crossing angle (error ~1e-8 rad near
//    asin(+-1)) cannot wrap through 0, and the bound a - beta needs no wrap test.
//  * eps covers the FP32 error of a and beta and the FP64 error of the exact
//    crossing angle.
// A disk whose bound reaches the current minimum travel angle cannot lower the
// minimum, so its exact GetFirstCrossTime can be skipped without changing any
// result.
template <bool CW>
__device__ __forceinline__ float d_travel_lb(const Orb& o, double px, double py, double d2)
{
    const double vx = px - o.cx, vy = py - o.cy;
    const double gx = o.u0x - vx, gy = o.u0y - vy;  // P0 - C
    const double gap = (double)o.r + 1e-6 * o.R;
    if (!(gx * gx + gy * gy >= gap * gap)) return 0.0f;
    const float ratio = o.r * rsqrtf((float)d2);
    if (!(ratio <= 0.5f)) return 0.0f;
    const double crs = o.u0x * vy - o.u0y * vx;
    const double dot = o.u0x * vx + o.u0y * vy;
    float a = atan2f((float)(CW ? -crs : crs), (float)dot);  // (-pi, pi]
    a = (a < 0.0f) ? a + 6.2831855f : a;
    const float lo = a - 1.0473f * ratio - (2e-6f + 1e-5f * a);
    return lo > 0.0f ? lo : 0.0f;
}

// Per-particle trace state and work counters.
struct TraceState {
    Isect nx;            // entry point of the next cell to visit
    double lifetime;     // running minimum (INF until a hit)
    int cells;           // cells visited
    int tested;          // impurities tested
    int edge_hits;       // accepted cell-edge intersections (Update_Best_Intersect calls)
    int cross_calls;     // GetFirstCrossTime calls of the reference (CirclesCross true)
    int cross_exact;     // of those, evaluated exactly here (the rest are screened out)
};

// Trace outcome flags.
#define ES_F_ESCAPED 1
#define ES_F_AT_BOUND 2
#define ES_F_CAP 4

__device__ __forceinline__ void d_trace_begin(TraceState& st, double px, double py, double phi0, int c0x, int c0y)
{
    st.nx.x = px;
    st.nx.y = py;
    st.nx.ang = phi0;
    st.nx.cx = c0x;
    st.nx.cy = c0y;
    st.lifetime = ES_INF;
    st.cells = st.tested = st.edge_hits = st.cross_calls = st.cross_exact = 0;
}

// One iteration of the TraceOrbit cell loop (lifetime.h:101-140).  Returns
// true when the trace has ended; then *flags holds ES_F_ESCAPED / ES_F_CAP
// and st.lifetime the raw lifetime before the bound (lifetime.h:142-147).
template <bool CW>
__device__ __forceinline__ bool d_step(const Orb& o, double phi0, const DevGrid& g, int cap, TraceState& st,
                                       int& flags)
{
    const Isect entry = st.nx;
    const bool avail = d_next_cell<CW>(o, phi0, g, entry, st.nx, st.edge_hits);
    const double rx = entry.ang;
    const double ry = avail ? st.nx.ang : phi0;
    const int ci = entry.cy * g.cpr + entry.cx;
    const int s = (ci > 0) ? g.cell_end[ci - 1] : 0;
    const int e = g.cell_end[ci];
    st.cells += 1;
    st.tested += (e - s);
#if ES_TRACE_SCREEN
    // The reference tests every impurity of the cell and keeps the smallest
    // GetFirstCrossTime (lifetime.h:117-123); a minimum does not depend on the
    // order, so: pass 1 runs CirclesCross (details.h:141-151) on all of them and
    // finds the crossing disk with the smallest screening bound; that one is
    // solved first, then pass 2 solves every other crossing disk whose bound
    // does not already exceed the current minimum.
    double lifetime = st.lifetime;
    float best_lb = 3.0e38f;
    int best_i = -1, ncross = 0;
    for (int i = s; i < e; i += 4) {
        double2 p[4];
#pragma unroll
        for (int u = 0; u < 4; u++) p[u] = (i + u < e) ? g.ixy[i + u] : make_double2(0.0, 0.0);
        int hits = 0;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const double qx = o.cx - p[u].x, qy = o.cy - p[u].y;
            const double d2 = qx * qx + qy * qy;
            hits |= ((i + u < e) && (d2 < o.radd2) && (d2 > o.rmin2)) ? (1 << u) : 0;
        }
#pragma unroll 1
        while (hits) {
            const int u = __builtin_ctz(hits);
            hits &= hits - 1;
            const double hx = (u == 0) ? p[0].x : (u == 1) ? p[1].x : (u == 2) ? p[2].x : p[3].x;
            const double hy = (u == 0) ? p[0].y : (u == 1) ? p[1].y : (u == 2) ? p[2].y : p[3].y;
            const double qx = o.cx - hx, qy = o.cy - hy;
            const float lb = d_travel_lb<CW>(o, hx, hy, qx * qx + qy * qy);
            ncross += 1;
            if (lb < best_lb) {
                best_lb = lb;
                best_i = i + u;
            }
        }
    }
    st.cross_calls += ncross;
    // Pass 2.  Its first round (i = s - 4) holds only the best-bound disk, which
    // is always solved; later rounds skip it.  One call site of the solver.
    const int last_i = (ncross > 1) ? e : s;
    for (int i = s - 4; best_i >= 0 && i < last_i; i += 4) {
        const bool first = (i < s);
        double2 p[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int k = first ? best_i : i + u;
            p[u] = (first ? (u == 0) : (k < e)) ? g.ixy[k] : make_double2(0.0, 0.0);
        }
        int hits = 0;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const double qx = o.cx - p[u].x, qy = o.cy - p[u].y;
            const double d2 = qx * qx + qy * qy;
            const bool cand = first ? (u == 0)
                                    : ((i + u < e) && (i + u != best_i) && (d2 < o.radd2) && (d2 > o.rmin2));
            hits |= cand ? (1 << u) : 0;
        }
#pragma unroll 1
        while (hits) {
            const int u = __builtin_ctz(hits);
            hits &= hits - 1;
            const double hx = (u == 0) ? p[0].x : (u == 1) ? p[1].x : (u == 2) ? p[2].x : p[3].x;
            const double hy = (u == 0) ? p[0].y : (u == 1) ? p[1].y : (u == 2) ? p[2].y : p[3].y;
            if (!first) {
                const double qx = o.cx - hx, qy = o.cy - hy;
                const float lb = d_travel_lb<CW>(o, hx, hy, qx * qx + qy * qy);
                if ((double)lb >= o.w * lifetime) continue;
            }
            const double t = d_first_cross_time<CW>(o, hx, hy, rx, ry);
            lifetime = (t < lifetime) ? t : lifetime;
            st.cross_exact += 1;
        }
    }
#else
    // Every impurity of the cell: CirclesCross (details.h:141-151), then
    // GetFirstCrossTime for the crossing ones in impurity order, keeping the
    // smallest time (lifetime.h:117-123).  Four impurities' loads in flight.
    double lifetime = st.lifetime;
    for (int i = s; i < e; i += 4) {
        double2 p[4];
#pragma unroll
        for (int u = 0; u < 4; u++) p[u] = (i + u < e) ? g.ixy[i + u] : make_double2(0.0, 0.0);
        int hits = 0;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const double qx = o.cx - p[u].x, qy = o.cy - p[u].y;
            const double d2 = qx * qx + qy * qy;
            hits |= ((i + u < e) && (d2 < o.radd2) && (d2 > o.rmin2)) ? (1 << u) : 0;
        }
#pragma unroll 1
        while (hits) {
            const int u = __builtin_ctz(hits);
            hits &= hits - 1;
            const double hx = (u == 0) ? p[0].x : (u == 1) ? p[1].x : (u == 2) ? p[2].x : p[3].x;
            const double hy = (u == 0) ? p[0].y : (u == 1) ? p[1].y : (u == 2) ? p[2].y : p[3].y;
            const double t = d_first_cross_time<CW>(o, hx, hy, rx, ry);
            lifetime = (t < lifetime) ? t : lifetime;
            st.cross_calls += 1;
            st.cross_exact += 1;
        }
    }
#endif
    st.lifetime = lifetime;
    if (lifetime < ES_INF) {
        flags = 0;
        return true;
    }
    if (!avail) {
        flags = ES_F_ESCAPED;
        st.lifetime = 0.0;
        return true;
    }
    if (st.cells >= cap) {
        flags = ES_F_CAP;
        st.lifetime = 0.0;
        return true;
    }
    return false;
}
