// es_trace32.h — one cell of a particle's trace in FP32, with flags for the
// cases FP32 cannot decide (the mixed-precision path; es_trace.h is FP64).
//
// Same trace as TraceOrbit (lifetime.h:78-148), in other coordinates.  Every
// length is taken relative to the particle's start point S and divided by the
// orbit radius R; the orbit is
//     X(theta) = n (1 - cos theta) + t sin theta,
// n the unit vector from S to the orbit centre, theta >= 0 the angle of the
// reference's travel (GetCrossAngle, details.h:111-116), t its direction at
// S — which is -(cos phi, sin phi) in both rotation senses: the reference's
// time runs along the orbit against the velocity it starts with.  With
// w = tan(theta / 2) a crossing with a disk or a cell-edge line is a
// quadratic A w^2 + 2 Bh w + C = 0, solved in the stable form
// s = -Bh - sign(Bh) sqrt(disc), roots s / A and C / s.  A root is kept as w
// when |w| <= 1, else as v = 1 / w, so nothing overflows; the key
//     k = w / (1 + w)        (0 <= w <= 1)      k = 1 / (1 + v)      (0 <= v <= 1)
//     k = 2 + w / (1 - w)    (-1 <= w < 0)      k = 2 + 1 / (v - 1)  (-1 <= v < 0)
// orders travel angles over one revolution (theta in [0, 2 pi) <-> k in [0, 2))
// with no trigonometric call; only the final hit is turned into an angle.
//
// A disk is screened in FP32; for the disks that may cross, the quantities
// that would cancel are formed from FP64 differences: q = (|P - S|^2 - r^2) / R^2
// (small when the start is near the disk), g = |P - C|^2 / R^2 - 1 and
// A = 4 (1 - b) + q; the discriminant is written as
// (g - (-2 rr + rr^2)) ((2 rr + rr^2) - g), rr = r / R — the product of the two
// CirclesCross margins (details.h:141-151), zero only at tangency.
//
// The reference's valid angle range starts the first cell at the raw particle
// angle phi (lifetime.h:101-106), which is negative for the incoherent
// window's first quadrant, and ends a cell without a successor at that raw
// angle too; AngleInRange (details.h:119-136) and the "next arc holds the
// start angle" test of GetNextCell (cell_grid.h:177-180) then accept other
// arcs than the one travelled.  With m = -phi > 0 these are, in travel angle:
//   first cell, exit at theta_e:  counter-clockwise: no next cell when
//     theta_e < m, else disks count on [m, theta_e); clockwise: (0, theta_e)
//     when theta_e < 2 pi - m, else none;
//   later cell entered at theta_i, no next cell:  counter-clockwise:
//     (theta_i, 2 pi) when theta_i >= m, else none; clockwise: (theta_i, 2 pi - m)
//     when theta_i < 2 pi - m, else (0, 2 pi - m) and (theta_i, 2 pi);
//   first cell without a next cell: the whole revolution.
// They are reproduced here, with m and 2 pi - m as keys.
//
// Cases FP32 cannot decide raise ES_F_RECOMPUTE and end the trace: the item is
// traced again in FP64.  They are: a disk within FP32 error of tangency to the
// orbit, an edge line within error of tangency, an exit within error of a
// cell corner (the corner snap of cell_grid.h:162-175) or of the entry point
// (DifferentPoint, :74-77), an edge point of the first cell within the FP64
// angle floor of the start (ahead or behind), a start point on a cell edge, a
// crossing within
// error of its cell's range ends, an exit or a hit within error of a full
// revolution or of the angles m, 2 pi - m above, a hit within the FP64
// path's angle floor of the start, a hit or an exit near the top or bottom
// of the orbit (where the reference's angle of it can be NaN), and a lifetime
// within 1e-5 of the bound time (lifetime.h:142-146).
#pragma once

#include "es_trace.h"

#define ES_F_RECOMPUTE 8

// One orbit in the start point's frame.
struct Orb32 {
    double sx, sy;          // start point (FP64)
    double inv_r;           // 1 / R
    double nxd, nyd;        // start -> centre, unit (FP64, for the crossing disks)
    double txd, tyd;        // direction of travel at the start (FP64)
    double r2;              // r^2
    float nx, ny, tx, ty;   // the same in FP32
    float rr, rr2;          // r / R, (r / R)^2
    float km, km2;          // keys of theta = m and 2 pi - m (quirk only)
    bool ccw;               // counter-clockwise orbit
    bool quirk;             // raw start angle phi < 0 (m = -phi)
};

// A root of the w-quadratic: z = w (inv == false, |z| <= 1) or z = 1 / w.
struct Root32 {
    float z;
    bool inv;
};

__device__ __forceinline__ float t32_key(Root32 r)
{
    if (!r.inv) return (r.z >= 0.0f) ? r.z / (1.0f + r.z) : 2.0f + r.z / (1.0f - r.z);
    return (r.z > 0.0f) ? 1.0f / (1.0f + r.z) : 2.0f + 1.0f / (r.z - 1.0f);
}

// Travel angle of a root: 2 atan(w) (+2 pi when w < 0), or pi - 2 atan(v).
__device__ __forceinline__ double t32_theta(Root32 r)
{
    const double a = 2.0 * (double)atanf(r.z);
    if (r.inv) return ES_PI - a;
    return (r.z >= 0.0f) ? a : ES_PI2 + a;
}

// Coordinate along (n_c, t_c) of the orbit point at root r:
// (2 n w^2 + 2 t w) / (1 + w^2)  or  (2 n + 2 t v) / (1 + v^2).
__device__ __forceinline__ float t32_coord(Root32 r, float n_c, float t_c)
{
    const float z2 = r.z * r.z;
    const float num = r.inv ? (2.0f * n_c + 2.0f * t_c * r.z) : (2.0f * n_c * z2 + 2.0f * t_c * r.z);
    return num / (1.0f + z2);
}

// Roots of A w^2 + 2 Bh w + C = 0 given disc = Bh^2 - A C >= 0.  Returns
// false when s = 0 (a double root at w = 0 or a degenerate quadratic).
__device__ __forceinline__ bool t32_roots(float A, float Bh, float C, float disc, Root32& r1, Root32& r2)
{
    const float h = sqrtf(disc > 0.0f ? disc : 0.0f);
    const float s = (Bh >= 0.0f) ? (-Bh - h) : (-Bh + h);
    if (s == 0.0f) return false;
    // root 1 = s / A, root 2 = C / s
    if (fabsf(s) <= fabsf(A)) {
        r1.z = s / A;
        r1.inv = false;
    } else {
        r1.z = A / s;
        r1.inv = true;
    }
    if (fabsf(C) <= fabsf(s)) {
        r2.z = C / s;
        r2.inv = false;
    } else {
        r2.z = s / C;
        r2.inv = true;
    }
    return true;
}

struct Trace32 {
    int cx, cy;             // current cell
    int in_edge;            // edge entered through (0 top, 1 bottom, 2 right, 3 left), -1 at the start
    float k_in;             // key of the entry point
    float best_k;           // key of the first hit so far (2 when none)
    Root32 best;            // its root
    int cells, tested, edge_hits, cross_calls;
};

__device__ __forceinline__ void t32_begin(Trace32& st, int c0x, int c0y)
{
    st.cx = c0x;
    st.cy = c0y;
    st.in_edge = -1;
    st.k_in = 0.0f;
    st.best_k = 2.0f;
    st.best.z = 0.0f;
    st.best.inv = false;
    st.cells = st.tested = st.edge_hits = st.cross_calls = 0;
}

// Error bands of the flags: relative, a few hundred FP32 ulps; absolute, in
// key units (half a small travel angle), the FP64 path's own floor — its
// angles come from asin(y / R), off by up to 1e-8 rad near y = +-R
// (details.h:93-107), so it may decide either way within that distance.
#define T32_REL 2e-5f
#define T32_ABS 1e-8f
// Where the reference compares two of its angles (a crossing against a range
// end, an edge point against the start), each may be off by ~3e-8 rad near
// y = +-R (asin of a value within a few ulps of 1): a band of 8e-8 rad in keys.
#define T32_END (4.0f * T32_ABS)

__device__ __forceinline__ bool t32_near(float a, float b, float band) { return fabsf(a - b) <= band; }

// A point of the orbit at the top or bottom (y - cy = +-R) can come out of the
// reference's FP64 operations a rounding step outside the circle: GetAngle
// then takes asin of a value above 1 (details.h:93-107) and the angle is NaN,
// so AngleInRange rejects that crossing (details.h:119-136) and GetNextCell
// never picks that exit (cell_grid.h:80-89).  The NaN needs the point within
// ~3e-8 rad of the pole; the FP32 coordinate below is good to ~7e-7 (in units
// of R), so a hit or an exit within T32_POLE of a pole goes to FP64.
#define T32_POLE 2e-6f

// x of (X - centre) / R for the orbit point at root r: (X - S) / R - n.
__device__ __forceinline__ bool t32_near_pole(const Orb32& o, Root32 r)
{
    return fabsf(t32_coord(r, o.nx, o.tx) - o.nx) < T32_POLE;
}

// Exit of the current cell and the reference's accepted travel angles in it.
struct Cell32 {
    float k_in, band_in;    // entry key and its error band
    float k_out, band_out;  // exit key (2 without an exit) and its band
    float a1, b1, a2, b2;   // accepted key ranges (a1, b1), (a2, b2)
    float b_hi;             // upper end of both ranges plus the band
    int e_out;              // exit edge (-1: none)
    int ncx, ncy;           // next cell
    bool avail;             // a next cell exists and the trace goes on there
};

// First half of one cell: the exit over the 4 edges and the key ranges in
// which disks count.  Returns false when FP32 cannot decide them.
__device__ __forceinline__ bool t32_exit(const Orb32& o, const DevGrid& g, Trace32& st, Cell32& c)
{
    const double L = g.L;
    const double x0 = g.start + (double)st.cx * L, y0 = g.start + (double)st.cy * L;
    // cell edges in the start frame, in units of R
    const float xl = (float)((x0 - o.sx) * o.inv_r), xh = (float)((x0 + L - o.sx) * o.inv_r);
    const float yl = (float)((y0 - o.sy) * o.inv_r), yh = (float)((y0 + L - o.sy) * o.inv_r);
    const float k_in = st.k_in;
    const bool first = st.in_edge < 0;
    const float band_in = T32_REL * k_in + T32_ABS;

    // ---- exit of the cell: smallest key after the entry over the 4 edges
    float k_out = 2.0f;
    int e_out = -1;
    Root32 r_out;           // the exit's root
    r_out.z = 0.0f;
    r_out.inv = false;
    float k_second = 2.0f;   // best key of another edge, for the corner test
    float k_flag = 3.0f;     // smallest key of an event FP32 cannot decide
    bool bad = false;
#pragma unroll 1
    for (int e = 0; e < 4; e++) {
        const bool vert = (e >= 2);                  // right / left edges are x = const
        const float cc = (e == 0) ? yh : (e == 1) ? yl : (e == 2) ? xh : xl;
        const float n_c = vert ? o.nx : o.ny, t_c = vert ? o.tx : o.ty;     // across the edge
        const float n_o = vert ? o.ny : o.nx, t_o = vert ? o.ty : o.tx;     // along the edge
        const float lo = vert ? yl : xl, hi = vert ? yh : xh;
        const float cband = T32_REL * (fabsf(lo) + fabsf(hi));
        // n_c (1 - cos) + t_c sin = cc  ->  (2 n_c - cc) w^2 + 2 t_c w - cc = 0
        const float A = 2.0f * n_c - cc;
        const float disc = t_c * t_c + cc * (2.0f * n_c - cc);
        const float dscale = t_c * t_c + fabsf(2.0f * n_c * cc) + cc * cc;
        if (disc <= T32_REL * dscale) {
            if (disc > -T32_REL * dscale) {
                // Near tangency: undecidable if a crossing could lie on the
                // edge.  disc = 1 - d^2 (d: the line's distance from the
                // centre, in units of R) is the squared half-chord, so any
                // crossing lies between the touching point and the roots of
                // the largest discriminant within the band — a half-chord of
                // up to sqrt(T32_REL dscale) ~ 5e-3 R, far wider than cband.
                Root32 tp;
                tp.inv = fabsf(t_c) > fabsf(A);
                tp.z = tp.inv ? (A / -t_c) : (-t_c / A);
                float amin = t32_coord(tp, n_o, t_o), amax = amin;
                float kmin = t32_key(tp), kmax = kmin;
                Root32 rw[2];
                if (t32_roots(A, t_c, -cc, T32_REL * dscale, rw[0], rw[1])) {
#pragma unroll
                    for (int k = 0; k < 2; k++) {
                        const float ak = t32_coord(rw[k], n_o, t_o), kk = t32_key(rw[k]);
                        amin = fminf(amin, ak);
                        amax = fmaxf(amax, ak);
                        kmin = fminf(kmin, kk);
                        kmax = fmaxf(kmax, kk);
                    }
                }
                if (amax > lo - cband && amin < hi + cband && kmax > k_in - band_in) k_flag = fminf(k_flag, kmin);
            }
            continue;
        }
        Root32 rt[2];
        if (!t32_roots(A, t_c, -cc, disc, rt[0], rt[1])) {
            bad = true;
            continue;
        }
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const float a = t32_coord(rt[k], n_o, t_o);
            if (!(a > lo - cband && a < hi + cband)) continue;
            const float kk = t32_key(rt[k]);
            // An edge point within the FP64 floor of the start, ahead or just
            // behind: the reference's angle of it (GetAngle, details.h:93-107,
            // off by up to ~2e-8 rad near y = +-R) may equal phi, which makes it
            // the exit (dphi = 0, cell_grid.h:80-89) and leaves no next cell
            // (AngleInRange of a zero-width range, cell_grid.h:177-180).
            if (first && (kk <= T32_END || kk >= 2.0f - T32_END)) {
                bad = true;
                continue;
            }
            // Near the entry the reference's two angles (each off by up to ~2e-8
            // rad near y = +-R) may order the points either way.
            const float band_e = band_in + T32_END;
            if (kk <= k_in - band_e) continue;  // behind the entry: after a full revolution
            if (a < lo + cband || a > hi - cband) {  // within error of a corner
                k_flag = fminf(k_flag, kk);
                continue;
            }
            st.edge_hits += 1;
            if (fabsf(kk - k_in) <= band_e) {
                if (e != st.in_edge) k_flag = fminf(k_flag, kk);  // another edge at the entry point
                continue;                                         // else: the entry point itself
            }
            if (kk < k_out) {
                if (e != e_out) k_second = k_out;
                k_out = kk;
                e_out = e;
                r_out = rt[k];
            } else if (e != e_out && kk < k_second) {
                k_second = kk;
            }
        }
    }
    const float band_out = T32_REL * k_out + T32_ABS;
    if (e_out >= 0 && k_second - k_out <= band_out) bad = true;  // two edges at one point
    if (k_flag <= k_out + band_out) bad = true;                 // an undecidable event before the exit
    if (e_out >= 0 && k_out >= 2.0f - 1e-5f) bad = true;        // exit near a full revolution
    if (e_out >= 0 && t32_near_pole(o, r_out)) bad = true;      // exit where the reference's angle may be NaN
    const int ncx = st.cx + ((e_out == 2) ? 1 : (e_out == 3) ? -1 : 0);
    const int ncy = st.cy + ((e_out == 0) ? 1 : (e_out == 1) ? -1 : 0);
    bool avail = e_out >= 0 && ncx >= 0 && ncx < g.cpr && ncy >= 0 && ncy < g.cpr;

    // ---- the reference's accepted travel angles in this cell: (a1, b1) and (a2, b2)
    float a1 = k_in, b1 = avail ? k_out : 2.0f, a2 = 2.0f, b2 = 2.0f;
    if (o.quirk) {
        const float bm = T32_REL * o.km + T32_END, bm2 = T32_REL * o.km2 + T32_END;
        if (first) {
            if (avail && o.ccw) {
                if (t32_near(k_out, o.km, bm)) bad = true;
                if (k_out < o.km) avail = false;  // the exit arc holds the raw start angle
            }
            if (!avail) {
                a1 = 0.0f;  // raw angle at both ends: the whole revolution
                b1 = 2.0f;
            } else if (o.ccw) {
                a1 = o.km;
            } else {
                if (t32_near(k_out, o.km2, bm2)) bad = true;
                if (k_out >= o.km2) a1 = b1 = 2.0f;
            }
        } else if (!avail) {
            if (o.ccw) {
                if (t32_near(k_in, o.km, bm)) bad = true;
                if (k_in < o.km) a1 = b1 = 2.0f;
            } else {
                if (t32_near(k_in, o.km2, bm2)) bad = true;
                if (k_in < o.km2) {
                    b1 = o.km2;
                } else {
                    a2 = k_in;
                    a1 = 0.0f;
                    b1 = o.km2;
                }
            }
        }
    }
    c.k_in = k_in;
    c.band_in = band_in;
    c.k_out = k_out;
    c.band_out = band_out;
    c.a1 = a1;
    c.b1 = b1;
    c.a2 = a2;
    c.b2 = b2;
    c.b_hi = fmaxf(b1, b2) + band_out;
    c.e_out = e_out;
    c.ncx = ncx;
    c.ncy = ncy;
    c.avail = avail;
    return !bad;
}

// What the FP64-formed disk test needs of the orbit and the cell.
struct Disk32Ctx {
    double sx, sy, inv_r, nxd, nyd, txd, tyd, r2;
    float glo, ghi, rr;
    float k_in, band_in, b_hi, a1, b1, a2, b2;
};

__device__ __forceinline__ Disk32Ctx t32_ctx(const Orb32& o, const Cell32& c)
{
    Disk32Ctx d;
    d.sx = o.sx;
    d.sy = o.sy;
    d.inv_r = o.inv_r;
    d.nxd = o.nxd;
    d.nyd = o.nyd;
    d.txd = o.txd;
    d.tyd = o.tyd;
    d.r2 = o.r2;
    d.glo = -2.0f * o.rr + o.rr2;
    d.ghi = 2.0f * o.rr + o.rr2;
    d.rr = o.rr;
    d.k_in = c.k_in;
    d.band_in = c.band_in;
    d.b_hi = c.b_hi;
    d.a1 = c.a1;
    d.b1 = c.b1;
    d.a2 = c.a2;
    d.b2 = c.b2;
    return d;
}

// FP32 screen of CirclesCross (details.h:141-151) for the disk at offset
// (dx, dy) from the start: |P - C|^2 / R^2 - 1 = |p|^2 - 2 p.n, with a band.
// False only for disks that certainly miss the orbit.
__device__ __forceinline__ bool t32_screen(const Orb32& o, float glo, float ghi, double dx, double dy)
{
    const float px = (float)(dx * o.inv_r), py = (float)(dy * o.inv_r);
    const float b0 = px * o.nx + py * o.ny;
    const float pp0 = px * px + py * py;
    const float g0 = pp0 - 2.0f * b0;
    const float band0 = 1e-5f * (pp0 + 2.0f * fabsf(b0)) + 1e-12f;
    return !(g0 < glo - band0 || g0 > ghi + band0);
}

// One disk that passed the screen, with the quantities that cancel formed
// from FP64 differences.  Returns false when FP32 cannot decide it.  Else:
// *cross = CirclesCross; for a crossing disk, its two roots rt[k] with keys
// kk[k], whether each lies in an accepted range (in[k]) and whether it lies
// within error of a range end (at_end[k]).
__device__ __forceinline__ bool t32_disk(const Disk32Ctx& c, double dx, double dy, bool& cross, Root32 rt[2],
                                         float kk[2], bool in[2], bool at_end[2])
{
    cross = false;
    const double ppn = (dx * dx + dy * dy) * (c.inv_r * c.inv_r);
    const double bn = (dx * c.nxd + dy * c.nyd) * c.inv_r;
    const float a = (float)((dx * c.txd + dy * c.tyd) * c.inv_r);  // along the travel
    const double qd = (dx * dx + dy * dy - c.r2) * (c.inv_r * c.inv_r);
    const float q = (float)qd;
    const float gg = (float)(ppn - 2.0 * bn);       // |P - C|^2 / R^2 - 1
    const float A = (float)((4.0 - 4.0 * bn) + qd);  // 4 (1 - b) + q (small near theta = pi)
    const float gband = T32_REL * (fabsf(gg) + 2.0f * c.rr) + 1e-14f * (float)(ppn + 2.0 * fabs(bn));
    if (fabsf(gg - c.glo) <= gband || fabsf(gg - c.ghi) <= gband) {
        // Within error of tangency: undecidable if a crossing could be in
        // range.  The crossings lie between the touching point and the roots
        // of the largest discriminant the band allows, (g - glo)(ghi - g) at
        // |g - glo| or |g - ghi| = gband; keys spanning the wrap at 0 / 2 are
        // undecidable too.
        Root32 tp;
        tp.inv = fabsf(2.0f * a) > fabsf(A);
        tp.z = tp.inv ? (A / (2.0f * a)) : (2.0f * a / A);
        float kmin = t32_key(tp), kmax = kmin;
        Root32 rw[2];
        if (t32_roots(A, -2.0f * a, q, gband * (c.ghi - c.glo + gband), rw[0], rw[1])) {
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const float kk = t32_key(rw[k]);
                kmin = fminf(kmin, kk);
                kmax = fmaxf(kmax, kk);
            }
        }
        if (kmax - kmin > 1.0f) return false;
        return !(kmax > c.k_in - c.band_in && kmin < c.b_hi);
    }
    if (!(gg > c.glo && gg < c.ghi)) return true;
    cross = true;
    if (!(q > 0.0f)) return false;  // the start lies outside every disk (k_positions)
    // |X - P|^2 = rr^2  ->  A w^2 - 4 a w + q = 0, disc = (2a)^2 - A q
    const float disc = (gg - c.glo) * (c.ghi - gg);
    if (!t32_roots(A, -2.0f * a, q, disc, rt[0], rt[1])) return false;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        kk[k] = t32_key(rt[k]);
        in[k] = (kk[k] > c.a1 && kk[k] < c.b1) || (kk[k] > c.a2 && kk[k] < c.b2);
        // A crossing within error of a range end (a cell boundary) may belong
        // to either cell, and FP32 keys can tie there.
        const float rb = 2e-6f * kk[k] + T32_END;
        at_end[k] = (c.a1 > 0.0f && t32_near(kk[k], c.a1, rb)) || (c.b1 < 2.0f && t32_near(kk[k], c.b1, rb)) ||
                    (c.a2 > 0.0f && c.a2 < 2.0f && t32_near(kk[k], c.a2, rb)) ||
                    (c.b2 < 2.0f && t32_near(kk[k], c.b2, rb));
        // First cell: a root just behind the start may come out just ahead of
        // it in the reference's angles (a hit at once): treat it as at the start.
        if (c.k_in == 0.0f && kk[k] >= 2.0f - T32_END) {
            at_end[k] = true;
            kk[k] = 0.0f;
        }
    }
    return true;
}

// Second half of one cell, after its disks: best_k / best the first hit in
// range (best_k = 2 without one), bad whether anything was undecidable.
// Returns true when the trace has ended (flags, theta as for t32_step).
__device__ __forceinline__ bool t32_finish(const Orb32& o, int cap, const Cell32& c, Trace32& st, bool bad,
                                           float best_k, Root32 best, int& flags, double& theta)
{
    st.best_k = best_k;
    st.best = best;
    if (!bad && o.quirk && best_k < 2.0f &&
        (t32_near(best_k, o.km, T32_REL * o.km + T32_END) || t32_near(best_k, o.km2, T32_REL * o.km2 + T32_END)))
        bad = true;  // a hit at the quirk angles
    if (!bad && best_k < T32_END) bad = true;  // a hit within the FP64 floor of the start
    if (!bad && best_k < 2.0f && t32_near_pole(o, best)) bad = true;  // a hit whose reference angle may be NaN
    if (bad) {
        flags = ES_F_RECOMPUTE;
        theta = 0.0;
        return true;
    }
    if (best_k < 2.0f) {
        flags = (best_k >= 2.0f - 1e-5f) ? ES_F_RECOMPUTE : 0;
        theta = t32_theta(best);
        return true;
    }
    if (!c.avail) {
        flags = ES_F_ESCAPED;
        theta = 0.0;
        return true;
    }
    if (st.cells >= cap) {
        flags = ES_F_CAP;
        theta = 0.0;
        return true;
    }
    st.cx = c.ncx;
    st.cy = c.ncy;
    const int e_out = c.e_out;
    st.in_edge = (e_out == 0) ? 1 : (e_out == 1) ? 0 : (e_out == 2) ? 3 : 2;  // the shared edge, seen from the new cell
    st.k_in = c.k_out;
    return false;
}

// One cell, one lane.  Returns true when the trace has ended; *flags then
// holds ES_F_ESCAPED / ES_F_CAP / ES_F_RECOMPUTE, and *theta the travel angle
// of the hit (0 without one).
__device__ __forceinline__ bool t32_step(const Orb32& o, const DevGrid& g, int cap, Trace32& st, int& flags,
                                         double& theta)
{
    Cell32 c;
    if (!t32_exit(o, g, st, c)) {
        flags = ES_F_RECOMPUTE;
        theta = 0.0;
        return true;
    }
    // ---- impurities of the cell: crossings with keys in the accepted ranges
    const int ci = st.cy * g.cpr + st.cx;
    const int s = (ci > 0) ? g.cell_end[ci - 1] : 0;
    const int en = g.cell_end[ci];
    st.cells += 1;
    st.tested += (en - s);
    const Disk32Ctx d = t32_ctx(o, c);
    bool bad = false;
    float best_k = st.best_k;
    Root32 best = st.best;
    for (int i = s; i < en; i += 4) {
        double2 p[4];
#pragma unroll
        for (int u = 0; u < 4; u++) p[u] = (i + u < en) ? g.ixy[i + u] : make_double2(0.0, 0.0);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (i + u >= en) continue;
            const double dx = p[u].x - o.sx, dy = p[u].y - o.sy;
            if (!t32_screen(o, d.glo, d.ghi, dx, dy)) continue;
            bool cross;
            Root32 rt[2];
            float kk[2];
            bool in[2], at_end[2];
            if (!t32_disk(d, dx, dy, cross, rt, kk, in, at_end)) {
                bad = true;
                st.cross_calls += cross ? 1 : 0;
                continue;
            }
            if (!cross) continue;
            st.cross_calls += 1;
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const float rb = 2e-6f * kk[k] + T32_END;
                if (at_end[k] && kk[k] < best_k + rb) bad = true;
                if (in[k] && kk[k] < best_k) {
                    best_k = kk[k];
                    best = rt[k];
                }
            }
        }
    }
    return t32_finish(o, cap, c, st, bad, best_k, best, flags, theta);
}
