/*
 * Test driver for the sanitizer build of the CPU oracle (oracle/es_oracle.c,
 * compiled into this file): bins generated fields (one with disks wider than a
 * cell, one with a single cell), then runs both passes at several fields with
 * the sigma, lifetime, temperature and image entry points on 1 and 3 threads,
 * printing every result as hex floats and integers.  The same driver built
 * without sanitizers must print the same text.
 */
#include <stdio.h>

#include "../../oracle/es_oracle.c"

static void print_metrics(const orc_metrics* m)
{
    printf("m %lld %lld %lld %lld %lld %lld %lld %lld %lld %lld\n", (long long)m->cells_passed,
           (long long)m->impurities_tested, (long long)m->particles_inside_impurity, (long long)m->particles_escaped,
           (long long)m->particles_at_bound, (long long)m->max_impurities_tested, (long long)m->max_cells_passed,
           (long long)m->trace_cap_hits, (long long)m->edge_hits, (long long)m->cross_calls);
}

static int run_case(uint64_t seed, double density, double radius, int target, int cpr, int ppr, int ppq, int threads)
{
    const double rs = 1.0e-6, ext = 2.5e-7;
    orc_grid* g = orc_grid_generate(seed, rs, ext, density, 0, radius, target, cpr);
    if (!g) return 1;
    int64_t ints[3];
    double dbl[3];
    orc_grid_info(g, ints, dbl);
    printf("grid cpr=%lld unique=%lld indexed=%lld cell=%a\n", (long long)ints[0], (long long)ints[1],
           (long long)ints[2], dbl[2]);
    double* xy = (double*)malloc((size_t)ints[2] * 2 * sizeof(double) + 16);
    int32_t* idx = (int32_t*)malloc((size_t)(ints[0] * ints[0]) * sizeof(int32_t));
    orc_grid_copy(g, xy, idx);
    double cs = 0;
    for (int64_t i = 0; i < 2 * ints[2]; i++) cs += xy[i];
    printf("csr end=%d sum=%a\n", idx[ints[0] * ints[0] - 1], cs);
    orc_sim_params p;
    memset(&p, 0, sizeof p);
    p.positions_per_row = ppr;
    p.particles_per_quadrant = ppq;
    p.region_size = rs;
    p.region_extends = ext;
    p.cell_size = dbl[2];
    p.tau = 1e-12;
    p.alpha = 0.3;
    p.is_clockwise = (int)(seed & 1);
    p.particle_speed = 7e5;
    p.nthreads = threads;
    const double fields[3] = {0.05, 2.0, 40.0};
    const double temps[3] = {2.0, 40.0, 100.0};
    const int64_t n = (int64_t)ppr * ppr * 4 * ppq;
    double* lts = (double*)malloc((size_t)n * sizeof(double));
    double* img = (double*)malloc((size_t)ppr * ppr * 3 * sizeof(double));
    for (int coherent = 0; coherent < 2; coherent++) {
        for (int b = 0; b < 3; b++) {
            orc_sigma s[3];
            orc_run_metrics rm;
            if (orc_compute_sigmas(g, &p, coherent, fields[b], temps, 3, s, &rm, lts)) return 1;
            print_metrics(&rm.pm);
            printf("real=%lld avg=%a\n", (long long)rm.real_particles, rm.avg_particle_lifetime);
            for (int t = 0; t < 3; t++) printf("s %a %a %a %a\n", s[t].xx_naive, s[t].xy_naive, s[t].xx_comp, s[t].xy_comp);
            orc_metrics m2;
            if (orc_compute_lifetimes(g, &p, coherent, fields[b], lts, &m2)) return 1;
            double lsum = 0;
            for (int64_t i = 0; i < n; i++) lsum += lts[i];
            printf("lt %a\n", lsum);
            orc_sigma s2[3];
            if (orc_derive_temperatures(g, &p, coherent, fields[b], lts, temps, 3, s2)) return 1;
            for (int t = 0; t < 3; t++) printf("d %a %a\n", s2[t].xx_comp, s2[t].xy_comp);
            if (orc_derive_images(g, &p, coherent, fields[b], lts, temps[1], img)) return 1;
            double isum = 0;
            for (int64_t i = 0; i < (int64_t)ppr * ppr * 3; i++) isum += img[i];
            printf("img %a\n", isum);
        }
    }
    free(img);
    free(lts);
    free(xy);
    free(idx);
    orc_grid_free(g);
    return 0;
}

int main(void)
{
    int bad = 0;
    bad |= run_case(101, 1.5e15, 1.1e-8, 4, 0, 9, 3, 1);   /* dense field, reference cell rule */
    bad |= run_case(202, 2.0e14, 6.0e-8, 4, 24, 7, 3, 3);  /* disks wider than a cell, 3 threads */
    bad |= run_case(303, 5.0e13, 2.0e-8, 4, 1, 5, 3, 1);   /* a single cell */
    printf("done %d\n", bad);
    return bad;
}
