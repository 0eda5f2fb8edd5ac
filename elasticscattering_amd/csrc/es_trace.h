// es_trace.h — device side of one particle's lifetime (gfx950, FP64).
//
// The floating-point operations and their order are those of the reference
// kernel library (src/sim/es/details.h, cell_grid.h, lifetime.h); the code is
// organised for the GPU: per-orbit constants are computed once per particle,
// the rotation sense is a template parameter, impurities are read from SoA
// arrays, and the start-cell "inside an impurity" test is hoisted out (it does
// not depend on phi or B and runs once per position in es_kernels.hip).
// Compiled with -ffp-contract=off so no multiply-add is fused.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "es_math.h"

#define ES_PI 3.141592653589793238463
#define ES_PI2 6.283185307179586
#define ES_HALF_PI 1.570796326794896619231
#define ES_INF 1.7976931348623157e308   /* DBL_MAX (constants.h:36) */

// Device view of the binned impurity field (Grid::GetSettings, Grid.h:82-91).
struct DevGrid {
    const double* ix;       // cell-ordered x of indexed impurities
    const double* iy;       // cell-ordered y
    const int32_t* cell_end;  // inclusive end offset per cell (Grid::imp_index)
    double start;           // spawn_region_start
    double size;            // spawn_region_size
    double L;               // cell_size
    double radius;          // impurity_radius
    int cpr;                // cells_per_row
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

// GetAngle + AngleVelocity (details.h:85-107).
template <bool CW>
__device__ __forceinline__ double d_angle(double x, double y, const Orb& o)
{
    const double y_dist = y - o.cy;
    const bool left_side = x < o.cx;
    double angle;
    if (fabs(y_dist) > o.eps_r) {
        angle = esm_asin(y_dist / o.R);
        angle = left_side ? ES_PI - angle : angle;
    } else {
        angle = left_side ? ES_PI : 0.0;
    }
    return d_smod(angle + (CW ? ES_HALF_PI : -ES_HALF_PI), ES_PI2);
}

// GetCrossAngle (details.h:111-116).
template <bool CW>
__device__ __forceinline__ double d_cross_angle(double p, double q)
{
    return d_smod(CW ? (p - q) : (q - p), ES_PI2);
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

// GetFirstCrossTime with GetCrossPoints inlined (details.h:153-206).
template <bool CW>
__device__ __forceinline__ double d_first_cross_time(const Orb& o, double px, double py, double rx, double ry)
{
    const double qx = o.cx - px, qy = o.cy - py;
    const double dist_squared = qx * qx + qy * qy;
    const double dist = sqrt(dist_squared);
    const double xs = (dist_squared + o.R2 - o.r2) / (2.0 * dist);
    const double ys = sqrt(o.R2 - xs * xs);
    const double ux = (px - o.cx) / dist, uy = (py - o.cy) / dist;
    const double bx = o.cx + ux * xs, by = o.cy + uy * xs;
    const double p1x = bx + uy * ys, p1y = by + (-ux) * ys;
    const double p2x = bx + (-uy) * ys, p2y = by + ux * ys;
    const double phi1 = d_angle<CW>(p1x, p1y, o);
    const double phi2 = d_angle<CW>(p2x, p2y, o);
    double tt = ES_INF;
    if (d_in_range<CW>(phi1, rx, ry)) {
        const double t = d_cross_angle<CW>(o.pa, phi1) / o.w;
        tt = (tt < t) ? tt : t;
    }
    if (d_in_range<CW>(phi2, rx, ry)) {
        const double t = d_cross_angle<CW>(o.pa, phi2) / o.w;
        tt = (tt < t) ? tt : t;
    }
    return tt;
}

// Update_Best_Intersect (cell_grid.h:80-89).
template <bool CW>
__device__ __forceinline__ void d_update_best(const Orb& o, double x, double y, int ncx, int ncy, double dp_eps,
                                              const Isect& last, Isect& best, double& best_dphi)
{
    const double ia = d_angle<CW>(x, y, o);
    const double dphi = d_cross_angle<CW>(last.ang, ia);
    if (dphi < best_dphi && (fabs(x - last.x) > dp_eps || fabs(y - last.y) > dp_eps)) {
        best.x = x;
        best.y = y;
        best.cx = ncx;
        best.cy = ncy;
        best.ang = ia;
        best_dphi = dphi;
    }
}

// UpdateFirstBoundaryIntersect (cell_grid.h:91-129) for the edge p1 -> p2.
template <bool CW>
__device__ __forceinline__ void d_edge(const Orb& o, int offx, int offy, double p1x, double p1y, double p2x,
                                       double p2y, double L, double dp_eps, const Isect& last, Isect& best,
                                       double& best_dphi, int& edge_hits)
{
    const double ux = (p2x - p1x) / L, uy = (p2y - p1y) / L;
    const double pd = ux * (o.cx - p1x) + uy * (o.cy - p1y);
    const double projx = p1x + ux * pd, projy = p1y + uy * pd;
    const double dx = projx - o.cx, dy = projy - o.cy;
    const double pcd = dx * dx + dy * dy;
    if (pcd >= o.R2) return;
    const double s = sqrt(o.R2 - pcd);
    const double tx = ux * s, ty = uy * s;
    const double i1x = projx + tx, i1y = projy + ty;
    const double i2x = projx - tx, i2y = projy - ty;
    const bool horizontal = fabs(ux) > fabs(uy);
    const double a1 = horizontal ? i1x : i1y, a2 = horizontal ? i2x : i2y;
    const double l0 = horizontal ? p1x : p1y, l1 = horizontal ? p2x : p2y;
    const bool v1 = (a1 > l0 && a1 < l1) || (a1 < l0 && a1 > l1);
    const bool v2 = (a2 > l0 && a2 < l1) || (a2 < l0 && a2 > l1);
    const int ncx = last.cx + offx, ncy = last.cy + offy;
    if (v1) d_update_best<CW>(o, i1x, i1y, ncx, ncy, dp_eps, last, best, best_dphi);
    if (v2) d_update_best<CW>(o, i2x, i2y, ncx, ncy, dp_eps, last, best, best_dphi);
    edge_hits += (int)v1 + (int)v2;
}

// GetNextCell (cell_grid.h:132-181).  `phi0` is the particle's start angle.
template <bool CW>
__device__ __forceinline__ bool d_next_cell(const Orb& o, double phi0, const DevGrid& g, const Isect& last, Isect& next,
                                            int& edge_hits)
{
    double best_dphi = ES_PI2;
    next = last;
    const double L = g.L;
    const double dp_eps = 1e-6 * L;
    const double llx = g.start + (double)last.cx * L;
    const double lly = g.start + (double)last.cy * L;
    const double lrx = llx + L, lry = lly + 0.0;
    const double trx = llx + L, try_ = lly + L;
    const double tlx = llx + 0.0, tly = lly + L;

    d_edge<CW>(o, 0, 1, tlx, tly, trx, try_, L, dp_eps, last, next, best_dphi, edge_hits);
    d_edge<CW>(o, 0, -1, llx, lly, lrx, lry, L, dp_eps, last, next, best_dphi, edge_hits);
    d_edge<CW>(o, 1, 0, lrx, lry, trx, try_, L, dp_eps, last, next, best_dphi, edge_hits);
    d_edge<CW>(o, -1, 0, llx, lly, tlx, tly, L, dp_eps, last, next, best_dphi, edge_hits);

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

struct TraceCounters {
    int64_t cells;
    int64_t tested;
    int escaped;
    int at_bound;
    int cap_hit;
    int edge_hits;     // accepted cell-edge intersections (Update_Best_Intersect calls)
    int cross_calls;   // GetFirstCrossTime calls (CirclesCross true)
};

// TraceOrbit (lifetime.h:78-148) after the start-cell test, from cell (c0x, c0y).
template <bool CW>
__device__ double d_trace(const Orb& o, double px, double py, double phi0, int c0x, int c0y, double bound_time,
                          const DevGrid& g, TraceCounters& tc)
{
    double lifetime = ES_INF;
    Isect nx;
    nx.x = px;
    nx.y = py;
    nx.ang = phi0;
    nx.cx = c0x;
    nx.cy = c0y;
    const int64_t cap = 4LL * g.cpr * g.cpr + 16;
    int64_t cells = 0, tested = 0;
    while (true) {
        const Isect entry = nx;
        const bool avail = d_next_cell<CW>(o, phi0, g, entry, nx, tc.edge_hits);
        const double rx = entry.ang;
        const double ry = avail ? nx.ang : phi0;
        const int ci = entry.cy * g.cpr + entry.cx;
        const int s = (ci > 0) ? g.cell_end[ci - 1] : 0;
        const int e = g.cell_end[ci];
        cells += 1;
        tested += (e - s);
        for (int i = s; i < e; i++) {
            const double ipx = g.ix[i], ipy = g.iy[i];
            const double qx = o.cx - ipx, qy = o.cy - ipy;
            const double d2 = qx * qx + qy * qy;
            if ((d2 < o.radd2) && (d2 > o.rmin2)) {
                const double t = d_first_cross_time<CW>(o, ipx, ipy, rx, ry);
                lifetime = (t < lifetime) ? t : lifetime;
                tc.cross_calls += 1;
            }
        }
        if (lifetime < ES_INF) break;
        if (!avail) {
            tc.escaped += 1;
            lifetime = 0.0;
            break;
        }
        if (cells >= cap) {
            tc.cap_hit += 1;
            lifetime = 0.0;
            break;
        }
    }
    tc.cells = cells;
    tc.tested = tested;
    if (lifetime > bound_time) {
        tc.at_bound += 1;
        return bound_time;
    }
    return lifetime;
}
