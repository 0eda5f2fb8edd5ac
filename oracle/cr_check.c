/*
 * cr_check.c — TEST INFRASTRUCTURE: checks that the elementary functions of
 * elasticscattering_amd/csrc/es_math.h are correctly rounded.
 *
 * The reference value of f(x) is GCC's libquadmath (binary128, 113-bit
 * significand, errors below one binary128 ulp) rounded to double; that is the
 * correctly rounded double unless f(x) lies within ~2^-110 (relative) of a
 * midpoint between two doubles, which random arguments do not hit.  For each
 * argument class the program counts:
 *   misround  — es_math.h results that differ from the reference (must be 0);
 *   glibc     — glibc results that differ from the reference (for the record);
 *   fast_fail — arguments whose fast phase fails the rounding test;
 *   fast_err  — the largest relative error of the fast phase's double-double
 *               (must stay below the rounding test's bound ESM_EPS_FAST).
 * Usage: cr_check N   (N arguments per class; prints one JSON object per class)
 */
#include <math.h>
#include <quadmath.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../elasticscattering_amd/csrc/es_math.h"

static uint64_t g_state = 0x9e3779b97f4a7c15ULL;

static double unif(void)  /* xorshift64*: uniform in [0, 1) */
{
    g_state ^= g_state >> 12;
    g_state ^= g_state << 25;
    g_state ^= g_state >> 27;
    return (double)((g_state * 0x2545F4914F6CDD1DULL) >> 11) * 0x1p-53;
}

enum { F_SIN, F_COS, F_ASIN, F_EXP };

/* Arguments of one class: uniform in [lo, hi], or log-spread (|x| = 2^-k (1 + u)). */
typedef struct {
    const char* name;
    int fn;
    double lo, hi;
    int kind;     /* 0 uniform, 1 log-spread magnitudes (signed), 2 1 - 2^-k u (signed), 3 near table points */
} Class;

static double draw(const Class* c)
{
    const double u = unif();
    if (c->kind == 1) {
        const double x = ldexp(1.0 + unif(), -(int)(u * 40.0));
        return unif() < 0.5 ? -x : x;
    }
    if (c->kind == 2) {
        const double x = 1.0 - ldexp(unif(), -(int)(u * 52.0));
        return unif() < 0.5 ? -x : x;
    }
    if (c->kind == 3) {  /* within a few ulps of a multiple of pi/4, pi/2 or of k/256 */
        const double base = (c->fn == F_EXP) ? floor(u * 256.0) * 0.693147180559945309 / 128.0
                                             : floor(u * 64.0) * 0.785398163397448310;
        double x = base;
        const int steps = (int)(unif() * 64.0) - 32;
        for (int k = 0; k < abs(steps); k++) x = nextafter(x, steps > 0 ? INFINITY : -INFINITY);
        return (c->fn == F_EXP) ? -x : x;
    }
    return c->lo + (c->hi - c->lo) * u;
}

static double esm(int fn, double x)
{
    switch (fn) {
    case F_SIN: return esm_sin(x);
    case F_COS: return esm_cos(x);
    case F_ASIN: return esm_asin(x);
    default: return esm_exp(x);
    }
}

static double glibc(int fn, double x)
{
    switch (fn) {
    case F_SIN: return sin(x);
    case F_COS: return cos(x);
    case F_ASIN: return asin(x);
    default: return exp(x);
    }
}

static __float128 quad(int fn, double x)
{
    switch (fn) {
    case F_SIN: return sinq((__float128)x);
    case F_COS: return cosq((__float128)x);
    case F_ASIN: return asinq((__float128)x);
    default: return expq((__float128)x);
    }
}

/* The fast phase's double-double and scale, as es_math.h evaluates it; returns 0 when it does not apply. */
static int fast_phase(int fn, double x, esm_dd* y, __float128* exact)
{
    const double ax = fabs(x);
    if (fn == F_ASIN) {
        if (!(ax < 1.0) || ax < 0x1p-26) return 0;
        const int big = ax > 0.5;
        double uh = ax, ul = 0.0;
        if (big) {
            const double z = (1.0 - ax) * 0.5;
            uh = sqrt(z);
            ul = fma(-uh, uh, z) / (uh + uh);
        }
        const esm_dd A = esm_asin_core(uh, ul);
        *y = big ? esm_pio2_minus_2(A) : A;
        *exact = asinq((__float128)ax);
        return 1;
    }
    if (fn == F_EXP) {
        if (!(x > -746.0) || x > 709.8 || ax < 0x1p-60) return 0;
        esm_dd r;
        int i, m;
        *y = esm_exp_fast(x, &r, &i, &m);
        *exact = expq((__float128)x) / ldexpq(1.0Q, m);
        return 1;
    }
    if (!(ax < ESM_TRIG_MAX) || ax < 0x1p-27) return 0;
    esm_dd r;
    const int q = esm_reduce_pio2(x, &r);
    const esm_sc v = esm_sincos_fast(r);
    const int use_sin = (fn == F_SIN) == ((q & 1) == 0);  /* which of sin r, cos r gives f(x) */
    *y = use_sin ? v.s : v.c;
    /* sin x = sin r, cos r, -sin r, -cos r and cos x = cos r, -sin r, -cos r, sin r for q = 0..3 */
    const __float128 f = quad(fn, x);
    *exact = ((fn == F_SIN) ? (q >= 2) : (q == 1 || q == 2)) ? -f : f;
    return 1;
}

int main(int argc, char** argv)
{
    const long n = argc > 1 ? atol(argv[1]) : 200000;
    const Class classes[] = {
        {"asin_uniform", F_ASIN, -1.0, 1.0, 0},
        {"asin_near_one", F_ASIN, 0, 0, 2},
        {"asin_small", F_ASIN, 0, 0, 1},
        {"sin_uniform", F_SIN, -16.0, 16.0, 0},
        {"sin_small", F_SIN, 0, 0, 1},
        {"sin_near_multiples", F_SIN, 0, 0, 3},
        {"sin_wide", F_SIN, -1e6, 1e6, 0},
        {"cos_uniform", F_COS, -16.0, 16.0, 0},
        {"cos_small", F_COS, 0, 0, 1},
        {"cos_near_multiples", F_COS, 0, 0, 3},
        {"exp_path", F_EXP, -20.0, 0.0, 0},
        {"exp_full", F_EXP, -745.2, 709.79, 0},
        {"exp_subnormal", F_EXP, -745.2, -707.0, 0},
        {"exp_small", F_EXP, 0, 0, 1},
        {"exp_near_table", F_EXP, 0, 0, 3},
    };
    for (size_t c = 0; c < sizeof(classes) / sizeof(classes[0]); c++) {
        const Class* k = &classes[c];
        long mis = 0, gmis = 0, ffail = 0;
        double worst = 0.0, first_bad = 0.0;
        for (long i = 0; i < n; i++) {
            const double x = draw(k);
            const double ref = (double)quad(k->fn, x);
            const double e = esm(k->fn, x);
            if (memcmp(&e, &ref, sizeof e) != 0 && !(e != e && ref != ref)) {
                if (mis == 0) first_bad = x;
                mis++;
            }
            if (glibc(k->fn, x) != ref) gmis++;
            esm_dd y;
            __float128 ex;
            if (fast_phase(k->fn, x, &y, &ex) && ex != 0) {
                const double err = fabs((double)((((__float128)y.h + (__float128)y.l) - ex) / ex));
                if (err > worst) worst = err;
                if (!esm_rn_ok(y.h, y.l, ESM_EPS_FAST)) ffail++;
            }
        }
        printf("{\"class\": \"%s\", \"n\": %ld, \"misround\": %ld, \"first_misround_arg\": \"%a\", \"glibc\": %ld, "
               "\"fast_fail\": %ld, \"fast_err_log2\": %.3f}\n",
               k->name, n, mis, first_bad, gmis, ffail, worst > 0 ? log2(worst) : -1000.0);
        fflush(stdout);
    }
    return 0;
}
