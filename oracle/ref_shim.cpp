// ref_shim.cpp — test infrastructure.  Compiles the reference's own kernel
// headers (src/sim/es/cell_grid.h, details.h, util.h) where they lie under the
// reference tree and exports flat C entry points with the same signatures as
// the oracle's test hooks (orc_*), so tests can compare the two builds value
// for value.  Built into oracle/_ref/ only where the reference tree exists.
// lifetime.h and Grid.cpp include <windows.h> and cannot be built here.
#include "src/sim/es/cell_grid.h"
#include "src/sim/es/util.h"

#define EXPORT extern "C" __attribute__((visibility("default")))

EXPORT double ref_smod(double a, double b) { return smod(a, b); }

EXPORT double ref_get_angle(double px, double py, double cx, double cy, double R, int cw)
{
    Orbit o(v2(cx, cy), R, cw != 0);
    return GetAngle(v2(px, py), &o);
}

EXPORT int ref_angle_in_range(double phi, double lo, double hi, int cw)
{
    return AngleInRange(phi, v2(lo, hi), cw != 0) ? 1 : 0;
}

EXPORT double ref_cross_angle(double p, double q, int cw) { return GetCrossAngle(p, q, cw != 0); }

EXPORT int ref_circles_cross(double cx, double cy, double R, double px, double py, double r)
{
    Orbit o(v2(cx, cy), R, false);
    return CirclesCross(&o, v2(px, py), r) ? 1 : 0;
}

EXPORT void ref_cross_points(double cx, double cy, double R, double px, double py, double r, double* out4)
{
    Orbit o(v2(cx, cy), R, false);
    v4 p = GetCrossPoints(&o, v2(px, py), r);
    out4[0] = p.x;
    out4[1] = p.y;
    out4[2] = p.z;
    out4[3] = p.w;
}

EXPORT double ref_first_cross_time(double cx, double cy, double R, int cw, double particle_angle,
                                   double ipx, double ipy, double ir, double w, double lo, double hi)
{
    Orbit o(v2(cx, cy), R, cw != 0);
    o.particle_angle = particle_angle;
    return GetFirstCrossTime(&o, v2(ipx, ipy), ir, w, v2(lo, hi));
}

EXPORT int ref_inside_impurity(double px, double py, double ix, double iy, double r)
{
    return InsideImpurity(v2(px, py), v2(ix, iy), r) ? 1 : 0;
}

EXPORT double ref_bound_time(double phi, double alpha, double w, int coherent, int is_electron, int is_future)
{
    return GetBoundTime(phi, alpha, w, coherent != 0, is_electron != 0, is_future != 0);
}

EXPORT double ref_bound_angle(double phi, double alpha, int cw) { return GetBoundAngle(phi, alpha, cw != 0); }

EXPORT void ref_orbit_center(double px, double py, double vx, double vy, double R, double vf, int is_electron, double* out2)
{
    v2 c = GetCyclotronOrbitCenter(v2(px, py), v2(vx, vy), R, vf, is_electron != 0);
    out2[0] = c.x;
    out2[1] = c.y;
}

EXPORT void ref_get_cell(double px, double py, double start, double size, int cpr, int* out2)
{
    v2i c = get_cell(v2(px, py), start, size, cpr);
    out2[0] = c.x;
    out2[1] = c.y;
}

EXPORT void ref_to_world(int cx, int cy, double start, double cell_size, double* out2)
{
    v2 w = to_world(v2i(cx, cy), start, cell_size);
    out2[0] = w.x;
    out2[1] = w.y;
}

EXPORT int ref_different_point(double ax, double ay, double bx, double by, double L)
{
    return DifferentPoint(v2(ax, ay), v2(bx, by), L) ? 1 : 0;
}

EXPORT int ref_within_bounds(int x, int y, int cpr) { return within_bounds(v2i(x, y), cpr) ? 1 : 0; }
EXPORT int ref_get_index(int x, int y, int cpr) { return get_index(v2i(x, y), cpr); }

EXPORT int ref_next_cell(double cx, double cy, double R, int cw, double phi,
                         double spawn_start, double spawn_size, double cell_size, int cpr,
                         double* last, int* last_cell, double* next, int* next_cell)
{
    Orbit o(v2(cx, cy), R, cw != 0);
    ImpuritySettings is;
    is.spawn_region_start = spawn_start;
    is.spawn_region_size = spawn_size;
    is.cell_size = cell_size;
    is.cells_per_row = cpr;
    is.impurity_radius = 0;
    Intersection a, n;
    a.position = v2(last[0], last[1]);
    a.incident_angle = last[2];
    a.dphi = last[3];
    a.entering_cell = v2i(last_cell[0], last_cell[1]);
    bool ok = GetNextCell(&o, phi, &is, &a, &n);
    last[3] = a.dphi;
    next[0] = n.position.x;
    next[1] = n.position.y;
    next[2] = n.incident_angle;
    next[3] = n.dphi;
    next_cell[0] = n.entering_cell.x;
    next_cell[1] = n.entering_cell.y;
    return ok ? 1 : 0;
}

EXPORT double ref_simpson_weight(int i, int dim) { return SimpsonWeight(i, dim); }
EXPORT double ref_get_sigma(double lt, double phi, double tau, double w) { return GetSigma(lt, phi, tau, w); }
