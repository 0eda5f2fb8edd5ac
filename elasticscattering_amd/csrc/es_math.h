/*
 * es_math.h — correctly rounded FP64 sin, cos, asin and exp for the
 * lifetime / sigma path.
 *
 * The same source is compiled for the gfx950 device (hipcc) and for the host
 * (g++ / gcc).  Every operation is an IEEE-754 basic operation (+ - * / sqrt),
 * an exact operation (fabs, rint, ldexp of a normal result, bit masking) or an
 * explicit fused multiply-add (fma(), correctly rounded on x86-64 and CDNA4),
 * and callers compile with -ffp-contract=off, so a result is the same on the
 * CPU and on the GPU.
 *
 * Why correctly rounded: the lifetime path branches on strict comparisons of
 * computed angles (AngleInRange, DifferentPoint, the corner snap), so a one-ulp
 * difference between two math libraries can flip a branch and change a
 * lifetime by O(1).  A correctly rounded function has one possible result, so
 * this path agrees with any other correctly rounded implementation (and with
 * glibc 2.35 wherever glibc rounds correctly: all but ~0.15 % of arguments).
 *
 * Method (Ziv's strategy; Ziv, "Fast evaluation of elementary mathematical
 * functions with correctly rounded last bit", ACM TOMS 17(3), 1991):
 *   1. a fast phase evaluates the function as a double-double yh + yl with a
 *      relative error below 2^-68 (table of expansion points + short
 *      polynomial, the leading terms in double-double arithmetic);
 *   2. a rounding test checks that every real value within that error of
 *      yh + yl rounds to yh (esm_rn_ok); it fails for about one argument in
 *      2^14;
 *   3. otherwise an accurate phase (relative error below 2^-100: Taylor series
 *      in double-double for sin, cos and exp; for asin one Newton step on the
 *      accurate sine) gives the result.  It is the correctly rounded value
 *      unless the exact value lies within 2^-100 (relative) of a midpoint
 *      between two doubles.
 * Double-double arithmetic: Dekker (1971) and Knuth's TwoSum, with TwoProduct
 * by fma.  Tables (es_math_tables.h) are values of the functions at table
 * points and Taylor coefficients, written by gen_es_math_tables.py from 300-bit
 * mpmath evaluations.
 *
 * Domains: asin: all doubles (NaN outside [-1, 1]); exp: all doubles
 * (overflow to +inf, subnormal results rounded correctly, which needs FP64
 * denormals, the gfx950 default); sin / cos: |x| < 2^20 (NaN beyond: the
 * argument reduction is a four-part Cody-Waite reduction by pi/2, exact in its
 * first step for |x| < 2^20; this path's arguments are below 16).
 */
#ifndef ES_MATH_H
#define ES_MATH_H

#include <stdint.h>
#if !defined(__HIPCC__)
#include <math.h>
#endif

#include "es_math_tables.h"

#if defined(__HIPCC__)
#define ESM_FN __host__ __device__ static inline
#define ESM_COLD __host__ __device__ static __attribute__((noinline))
#else
#define ESM_FN static inline
#define ESM_COLD static __attribute__((noinline))
#endif

/* The tables: device code reads a __constant__ copy, host code its own. */
#if defined(__HIP_DEVICE_COMPILE__)
#define ESM_TABLE static __constant__ const
#else
#define ESM_TABLE static const
#endif

ESM_TABLE double esm_asin_tab[ESM_ASIN_N * 16] = ESM_ASIN_TAB_INIT;
ESM_TABLE double esm_sc_tab[202 * 4] = ESM_SC_TAB_INIT;
ESM_TABLE double esm_exp_tab[128 * 2] = ESM_EXP_TAB_INIT;
ESM_TABLE double esm_invfact[32] = ESM_INVFACT_INIT;

#define ESM_EPS_FAST 0x1p-67   /* bound used by the rounding test (twice the analysed error) */
#define ESM_TRIG_MAX 0x1p20

ESM_FN uint64_t esm_bits(double x)
{
    union { double d; uint64_t u; } v;
    v.d = x;
    return v.u;
}

ESM_FN double esm_from_bits(uint64_t u)
{
    union { double d; uint64_t u; } v;
    v.u = u;
    return v.d;
}

/* ------------------------------------------------------------ double-double */

typedef struct { double h, l; } esm_dd;
typedef struct { esm_dd s, c; } esm_sc;

ESM_FN esm_dd esm_dd_make(double h, double l)
{
    esm_dd r;
    r.h = h;
    r.l = l;
    return r;
}

/* a + b = s + e exactly (Knuth). */
ESM_FN esm_dd esm_two_sum(double a, double b)
{
    const double s = a + b;
    const double bb = s - a;
    return esm_dd_make(s, (a - (s - bb)) + (b - bb));
}

/* a + b = s + e exactly when a = 0 or exponent(a) >= exponent(b) (Dekker). */
ESM_FN esm_dd esm_fast_two_sum(double a, double b)
{
    const double s = a + b;
    return esm_dd_make(s, b - (s - a));
}

/* a * b = p + e exactly (no overflow or underflow). */
ESM_FN esm_dd esm_two_prod(double a, double b)
{
    const double p = a * b;
    return esm_dd_make(p, fma(a, b, -p));
}

/* (a.h + a.l)(b.h + b.l), relative error below 2^-104. */
ESM_FN esm_dd esm_dd_mul(esm_dd a, esm_dd b)
{
    esm_dd p = esm_two_prod(a.h, b.h);
    p.l = fma(a.h, b.l, fma(a.l, b.h, p.l));
    return esm_fast_two_sum(p.h, p.l);
}

/* (a.h + a.l) + (b.h + b.l), both parts summed exactly before renormalising. */
ESM_FN esm_dd esm_dd_add(esm_dd a, esm_dd b)
{
    esm_dd s = esm_two_sum(a.h, b.h);
    const esm_dd t = esm_two_sum(a.l, b.l);
    s.l += t.h;
    s = esm_fast_two_sum(s.h, s.l);
    s.l += t.l;
    return esm_fast_two_sum(s.h, s.l);
}

ESM_FN esm_dd esm_dd_neg(esm_dd a)
{
    return esm_dd_make(-a.h, -a.l);
}

/*
 * Rounding test: 1 when yh is the round-to-nearest value of every real y with
 * |y - (yh + yl)| <= eps |yh|.  Needs yh = RN(yh + yl) (|yl| <= ulp(yh) / 2)
 * and 2^-968 <= |yh| < inf.  The nearest rounding boundary on yl's side is half
 * an ulp of yh away, a quarter of an ulp below a power of two.  eps is twice
 * the analysed error, which also covers the rounding of the left-hand side.
 */
ESM_FN int esm_rn_ok(double yh, double yl, double eps)
{
    const uint64_t b = esm_bits(yh);
    double half = esm_from_bits((b & 0x7ff0000000000000ULL) - 0x0350000000000000ULL); /* 2^(E - 53) */
    const int toward_zero = (yl < 0.0) != (yh < 0.0);
    if ((b & 0x000fffffffffffffULL) == 0 && toward_zero) half *= 0.5;
    return fabs(yl) + eps * fabs(yh) < half;
}

/* -------------------------------------------------------------- sin / cos */

/*
 * x = n pi/2 + r with |r| <= pi/4 (1 + 2^-50); returns n mod 4 and r as a
 * double-double.  |x| < 2^20: the first step is exact (x and n P1 are
 * multiples of 2^-53 and their difference is below 1), n P2 and n P3 are
 * exact products, and the rest carries an absolute error near 2^-140.
 */
ESM_FN int esm_reduce_pio2(double x, esm_dd* r)
{
    const double n = rint(x * ESM_INVPIO2);
    const double a = fma(-n, ESM_PIO2_1, x);
    const esm_dd p2 = esm_two_prod(n, ESM_PIO2_2);
    const esm_dd p3 = esm_two_prod(n, ESM_PIO2_3);
    const esm_dd s = esm_two_sum(a, -p2.h);
    const double t = (((s.l - p2.l) - p3.h) - p3.l) - n * ESM_PIO2_4;
    *r = esm_fast_two_sum(s.h, t);
    return (int)((int64_t)n & 3);
}

/*
 * sin and cos of r = rh + rl, |r| <= 0.79, |rl| <= ulp(rh)/2, as
 * double-doubles with relative error below 2^-69:  r = j/256 + rho, |rho| <=
 * 2^-9, sin r = S + C rho + (C (sin rho - rho) + S (cos rho - 1)) and
 * cos r = C - S rho + (C (cos rho - 1) - S (sin rho - rho)) with S, C =
 * sin, cos (j/256) as double-doubles; the brackets (at most 2^-18 of the
 * result) are polynomials evaluated in double.
 */
ESM_FN esm_sc esm_sincos_fast(esm_dd r)
{
    const int neg = r.h < 0.0;
    const double ah = neg ? -r.h : r.h;
    const double al = neg ? -r.l : r.l;
    const int j = (int)rint(ah * 256.0);
    const double ph = ah - (double)j * 0x1p-8;   /* exact */
    const double* T = esm_sc_tab + 4 * j;
    const double z = fma(ph, ph, 2.0 * ph * al); /* rho^2 */
    const double ps = fma(z, fma(z, -0x1.a01a01a01a01ap-13, 0x1.1111111111111p-7), -0x1.5555555555555p-3);
    const double pc = fma(z, fma(z, -0x1.6c16c16c16c17p-10, 0x1.5555555555555p-5), -0.5);
    const double sr = ph * z * ps;  /* sin rho - rho */
    const double cr = z * pc;       /* cos rho - 1 */
    esm_dd cp = esm_two_prod(T[2], ph);
    cp.l = fma(T[2], al, fma(T[3], ph, cp.l));
    esm_dd s = esm_two_sum(T[0], cp.h);
    s.l += (T[1] + cp.l) + fma(T[2], sr, T[0] * cr);
    s = esm_fast_two_sum(s.h, s.l);
    esm_dd sp = esm_two_prod(T[0], ph);
    sp.l = fma(T[0], al, fma(T[1], ph, sp.l));
    esm_dd c = esm_two_sum(T[2], -sp.h);
    c.l += (T[3] - sp.l) + fma(T[2], cr, -(T[0] * sr));
    c = esm_fast_two_sum(c.h, c.l);
    esm_sc out;
    out.s = neg ? esm_dd_neg(s) : s;
    out.c = c;
    return out;
}

/*
 * The accurate phase: sin and cos of r = rh + rl, |r| <= 0.79, with relative
 * error below 2^-102: the same table split, with sin rho and cos rho as
 * Taylor series to rho^11 and rho^12 in double-double arithmetic.
 */
ESM_COLD esm_sc esm_sincos_acc(double rh, double rl)
{
    const int neg = rh < 0.0;
    const double ah = neg ? -rh : rh;
    const double al = neg ? -rl : rl;
    const int j = (int)rint(ah * 256.0);
    const esm_dd p = esm_two_sum(ah - (double)j * 0x1p-8, al);
    const esm_dd z = esm_dd_mul(p, p);
    /* sin rho / rho = sum_k (-1)^k z^k / (2k+1)!, k <= 5 */
    esm_dd ps = esm_dd_make(-esm_invfact[22], -esm_invfact[23]);
    for (int k = 4; k >= 0; k--) {
        const double sg = (k & 1) ? -1.0 : 1.0;
        ps = esm_dd_add(esm_dd_make(sg * esm_invfact[2 * (2 * k + 1)], sg * esm_invfact[2 * (2 * k + 1) + 1]),
                        esm_dd_mul(z, ps));
    }
    const esm_dd sr = esm_dd_mul(p, ps);
    /* cos rho = sum_k (-1)^k z^k / (2k)!, k <= 6 */
    esm_dd pc = esm_dd_make(esm_invfact[24], esm_invfact[25]);
    for (int k = 5; k >= 0; k--) {
        const double sg = (k & 1) ? -1.0 : 1.0;
        pc = esm_dd_add(esm_dd_make(sg * esm_invfact[2 * (2 * k)], sg * esm_invfact[2 * (2 * k) + 1]),
                        esm_dd_mul(z, pc));
    }
    const double* T = esm_sc_tab + 4 * j;
    const esm_dd S = esm_dd_make(T[0], T[1]);
    const esm_dd C = esm_dd_make(T[2], T[3]);
    esm_sc out;
    const esm_dd s = esm_dd_add(esm_dd_mul(S, pc), esm_dd_mul(C, sr));
    out.s = neg ? esm_dd_neg(s) : s;
    out.c = esm_dd_add(esm_dd_mul(C, pc), esm_dd_neg(esm_dd_mul(S, sr)));
    return out;
}

/* sin(x) and cos(x) of one argument, each correctly rounded. */
ESM_FN void esm_sincos(double x, double* sp, double* cp)
{
    const double ax = fabs(x);
    if (!(ax < ESM_TRIG_MAX)) {
        *sp = x - x;  /* NaN for inf, NaN and |x| >= 2^20 */
        *cp = x - x;
        return;
    }
    if (ax < 0x1p-27) {  /* |x^3/6| and x^2/2 are below a quarter ulp */
        *sp = x;
        *cp = 1.0;
        return;
    }
    esm_dd r;
    const int q = esm_reduce_pio2(x, &r);
    esm_sc v = esm_sincos_fast(r);
    if (!esm_rn_ok(v.s.h, v.s.l, ESM_EPS_FAST) || !esm_rn_ok(v.c.h, v.c.l, ESM_EPS_FAST))
        v = esm_sincos_acc(r.h, r.l);
    const double s = v.s.h, c = v.c.h;
    switch (q) {
    case 0: *sp = s;  *cp = c;  break;
    case 1: *sp = c;  *cp = -s; break;
    case 2: *sp = -s; *cp = -c; break;
    default: *sp = -c; *cp = s; break;
    }
}

ESM_FN double esm_sin(double x)
{
    double s, c;
    esm_sincos(x, &s, &c);
    return s;
}

ESM_FN double esm_cos(double x)
{
    double s, c;
    esm_sincos(x, &s, &c);
    return c;
}

/* ------------------------------------------------------------------- asin */

/*
 * asin(u), u = uh + ul in [0, 1/2], |ul| <= 2^-53 uh, as a double-double with
 * relative error below 2^-69: u lies in the interval i = floor(256 u) of the
 * table, whose Taylor polynomial of degree 9 about its centre c (0 for i = 0)
 * has truncation error below 2^-85; h = uh - c is exact (Sterbenz), a0 + a1 h
 * is formed in double-double, h^2 (a2 + a3 h + ... + a9 h^7) (at most 2^-18 of
 * the result) in double.
 */
ESM_FN esm_dd esm_asin_core(double uh, double ul)
{
    int i = (int)(uh * 256.0);
    i = i > ESM_ASIN_N - 1 ? ESM_ASIN_N - 1 : i;
    const double c = i ? ((double)i + 0.5) * 0x1p-8 : 0.0;
    const double h = uh - c;
    const double* a = esm_asin_tab + 16 * i;
    double t = fma(a[11], h, a[10]);
    t = fma(t, h, a[9]);
    t = fma(t, h, a[8]);
    t = fma(t, h, a[7]);
    t = fma(t, h, a[6]);
    t = fma(t, h, a[5]);
    t = fma(t, h, a[4]);
    const double h2 = fma(h, h, 2.0 * h * ul);
    esm_dd p = esm_two_prod(a[2], h);
    p.l = fma(a[2], ul, fma(a[3], h, p.l));
    esm_dd s = esm_two_sum(a[0], p.h);
    s.l += (a[1] + p.l) + h2 * t;
    return esm_fast_two_sum(s.h, s.l);
}

/* pi/2 - 2 (a.h + a.l) for a in [0, pi/6], as a double-double. */
ESM_FN esm_dd esm_pio2_minus_2(esm_dd a)
{
    esm_dd r = esm_two_sum(ESM_PIO2_H, -2.0 * a.h);
    r.l += ESM_PIO2_L - 2.0 * a.l;
    return esm_fast_two_sum(r.h, r.l);
}

/*
 * The accurate phase of asin: A = asin(u) to 2^-69 is refined by one Newton
 * step on sin, A + (u - sin A) / cos A, with sin A from the accurate sine
 * (2^-102); the step's second-order error is below 2^-136.
 */
ESM_COLD double esm_asin_acc(int big, double ax, double Ah, double Al)
{
    double uh = ax, ul = 0.0;
    if (big) {  /* sqrt(z) as RN(sqrt z) + the exact residual over 2 RN(sqrt z) */
        const double z = (1.0 - ax) * 0.5;
        uh = sqrt(z);
        ul = fma(-uh, uh, z) / (uh + uh);
    }
    const esm_sc v = esm_sincos_acc(Ah, Al);
    const double d = ((uh - v.s.h) + (ul - v.s.l)) / v.c.h;  /* uh - sin A is exact (Sterbenz) */
    esm_dd A = esm_fast_two_sum(Ah, Al + d);
    if (big) A = esm_pio2_minus_2(A);
    return A.h;
}

/*
 * asin(x): |x| <= 1/2 directly; 1/2 < |x| < 1 as pi/2 - 2 asin(sqrt(z)),
 * z = (1 - |x|)/2 (exact), with sqrt(z) as a double-double.
 */
ESM_FN double esm_asin(double x)
{
    const double ax = fabs(x);
    if (!(ax < 1.0)) {
        if (ax == 1.0) return x > 0.0 ? ESM_PIO2_H : -ESM_PIO2_H;
        return (x - x) / (x - x);  /* |x| > 1 or NaN: NaN */
    }
    if (ax < 0x1p-26) return x;    /* x^3/6 is below a quarter ulp of x */
    const int big = ax > 0.5;
    double uh = ax, ul = 0.0;
    if (big) {
        const double z = (1.0 - ax) * 0.5;  /* exact, in [2^-54, 1/4] */
#ifdef __HIP_DEVICE_COMPILE__
        /* uh within an ulp of sqrt(z) by Newton steps from v_rsq_f64 (z is normal, no scaling),
           h within 2^-50 of 1/(2 uh); ul = (z - uh^2) h is the residual to 2^-50 of itself */
        const double y = __builtin_amdgcn_rsq(z);
        double g = z * y, h = 0.5 * y;
        const double e = fma(-h, g, 0.5);
        g = fma(g, e, g);
        h = fma(h, e, h);
        g = fma(fma(-g, g, z), h, g);
        uh = fma(fma(-g, g, z), h, g);
        ul = fma(-uh, uh, z) * h;
#else
        uh = sqrt(z);
        ul = fma(-uh, uh, z) / (uh + uh);
#endif
    }
    const esm_dd A = esm_asin_core(uh, ul);
    const esm_dd R = big ? esm_pio2_minus_2(A) : A;
    double r = R.h;
    if (!esm_rn_ok(R.h, R.l, ESM_EPS_FAST)) r = esm_asin_acc(big, ax, A.h, A.l);
    return x < 0.0 ? -r : r;
}

/* -------------------------------------------------------------------- exp */

/*
 * exp(r) - 1 for |r| <= ln2/256 (1 + 2^-40), r = rh + rl, in the accurate
 * phase: the Taylor series to r^12 in double-double (truncation 2^-130).
 */
ESM_COLD esm_dd esm_exp_acc(double rh, double rl, int i)
{
    const esm_dd r = esm_dd_make(rh, rl);
    esm_dd s = esm_dd_make(esm_invfact[24], esm_invfact[25]);
    for (int k = 11; k >= 1; k--) s = esm_dd_add(esm_dd_make(esm_invfact[2 * k], esm_invfact[2 * k + 1]), esm_dd_mul(r, s));
    const esm_dd e = esm_dd_add(esm_dd_make(1.0, 0.0), esm_dd_mul(r, s));
    return esm_dd_mul(esm_dd_make(esm_exp_tab[2 * i], esm_exp_tab[2 * i + 1]), e);
}

/*
 * A result below 2^-1021 (m <= -1022), rounded on the grid of 2^-1074:
 * U = y 2^(m + 1074) is at most 2^53, its nearest integer n is taken from the
 * double-double; fast = 1 asks whether the fast phase's error can move U across
 * a half-integer (then 0 is returned through *ok).
 */
ESM_FN double esm_exp_tiny(esm_dd y, int m, int fast, int* ok)
{
    const double uh = ldexp(y.h, m + 1074), ul = ldexp(y.l, m + 1074);
    const double n = rint(uh);
    const double d = (uh - n) + ul;   /* uh - n is exact */
    if (fast) {
        *ok = 0.5 - fabs(d) > ESM_EPS_FAST * uh + 0x1p-52;
        return n * 0x1p-1074;
    }
    *ok = 1;
    const double adj = (d > 0.5) ? 1.0 : (d < -0.5) ? -1.0 : 0.0;
    return (n + adj) * 0x1p-1074;
}

/*
 * The fast phase of exp for -746 < x <= 709.8, |x| >= 2^-60:
 * exp(x) = 2^m 2^(i/128) exp(r), x = (128 m + i) ln2/128 + r, |r| <= ln2/256,
 * r from a three-part Cody-Waite reduction (exact first step), exp(r) - 1 by
 * its Taylor polynomial to r^6 (truncation 2^-71, 1 + r carried in
 * double-double), the table value 2^(i/128) as a double-double.  Returns
 * 2^(i/128) exp(r) (relative error below 2^-69) and sets *r, *i, *m.
 */
ESM_FN esm_dd esm_exp_fast(double x, esm_dd* rr, int* ii, int* mm)
{
    const double k = rint(x * ESM_INV_LN2_128);
    const double a = fma(-k, ESM_LN2_128_1, x);          /* exact */
    const esm_dd p = esm_two_prod(k, ESM_LN2_128_2);
    const esm_dd s = esm_two_sum(a, -p.h);
    const esm_dd r = esm_fast_two_sum(s.h, (s.l - p.l) - k * ESM_LN2_128_3);
    const int ki = (int)k;
    const int i = ki & 127;
    *rr = r;
    *ii = i;
    *mm = (ki - i) / 128;
    const double rh = r.h;
    const double q = fma(rh, fma(rh, fma(rh, 0x1.6c16c16c16c17p-10, 0x1.1111111111111p-7), 0x1.5555555555555p-5),
                         0x1.5555555555555p-3);
    const double rh2 = rh * rh;
    esm_dd e = esm_fast_two_sum(1.0, rh);
    e.l += r.l + fma(0.5, rh2, rh2 * rh * q);
    const double* T = esm_exp_tab + 2 * i;
    esm_dd y = esm_two_prod(T[0], e.h);
    y.l = fma(T[0], e.l, fma(T[1], e.h, y.l));
    return esm_fast_two_sum(y.h, y.l);
}

ESM_FN double esm_exp(double x)
{
    if (!(x > -746.0)) return (x != x) ? x + x : 0.0;   /* NaN; exp(x) < 2^-1076 rounds to +0 */
    if (x > 709.8) return 0x1p1023 * 2.0;                /* overflow: +inf */
    if (fabs(x) < 0x1p-60) return 1.0 + x;
    esm_dd r;
    int i, m;
    esm_dd y = esm_exp_fast(x, &r, &i, &m);
    if (m >= -1021) {
        if (!esm_rn_ok(y.h, y.l, ESM_EPS_FAST)) y = esm_exp_acc(r.h, r.l, i);
        return ldexp(y.h, m);
    }
    int ok;
    const double res = esm_exp_tiny(y, m, 1, &ok);
    if (ok) return res;
    return esm_exp_tiny(esm_exp_acc(r.h, r.l, i), m, 0, &ok);
}

/*
 * a / b given y = RN(1/b), in three operations:  q0 = RN(a * y),
 * r = a - b * q0 (exact with one fma), RN(q0 + r * y).  By Markstein's
 * theorem (IA-64 division; Muller et al., Handbook of Floating-Point
 * Arithmetic, 2nd ed., sec. 5.3) this is the correctly rounded quotient,
 * i.e. bit-identical to IEEE a / b, whenever y is the correctly rounded
 * reciprocal and nothing overflows or underflows.  The kernels use it where
 * the divisor is fixed for a particle or a launch (y computed once with an
 * IEEE division); the CPU restatement divides, and the bit-exact lifetime
 * tests compare the two.
 */
ESM_FN double esm_div_rcp(double a, double b, double y)
{
    const double q0 = a * y;
    const double r = fma(-b, q0, a);
    return fma(r, y, q0);
}

#endif /* ES_MATH_H */
