/*
 * es_math.h — FP64 elementary functions used on the lifetime / sigma path.
 *
 * The same source is compiled for the gfx950 device (hipcc) and for the host
 * (g++ / gcc).  Every operation is an IEEE-754 basic operation (+ - * / sqrt),
 * an exact operation (fabs, floor, rint, ldexp, bit masking) or an explicit
 * fused multiply-add (fma(), correctly rounded on both x86-64 and CDNA4), and
 * callers compile with -ffp-contract=off.  The results are therefore
 * bit-identical on the CPU and on the GPU, which is what lets the device
 * lifetimes be compared bit for bit with a CPU restatement: the lifetime path
 * branches on strict inequalities of computed angles, and a one-ulp
 * difference between two libm implementations can flip one of them.
 *
 * Method: fdlibm 5.3 (Sun Microsystems, 1993), whose notice follows —
 *   "Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.
 *    Developed at SunSoft, a Sun Microsystems, Inc. business.
 *    Permission to use, copy, modify, and distribute this software is freely
 *    granted, provided that this notice is preserved."
 * The minimax coefficients are fdlibm's (k_sin.c, k_cos.c, e_asin.c,
 * e_exp.c).  The argument reduction for sin/cos is a three-part Cody-Waite
 * reduction by pi/2 written with fma (pi/2 = P1 + P2 + P3 to ~160 bits);
 * it is exact in its first step and valid for |x| < 2^30, far beyond the
 * |x| < 16 this path produces (phi in [-pi/4, 2pi], plus omega*lifetime,
 * which is below one revolution).
 *
 * Accuracy against glibc 2.35 is measured by the test suite (max |ulp|).
 */
#ifndef ES_MATH_H
#define ES_MATH_H

#include <stdint.h>
#if !defined(__HIPCC__)
#include <math.h>
#endif

#if defined(__HIPCC__)
#define ESM_FN __host__ __device__ static inline
#else
#define ESM_FN static inline
#endif

/* pi/2 split into three doubles: P1 (53 bits), P2, P3. */
#define ESM_INVPIO2 6.36619772367581382433e-01   /* 0x3FE45F306DC9C883 */
#define ESM_PIO2_1  1.57079632679489655800e+00   /* 0x3FF921FB54442D18 */
#define ESM_PIO2_2  6.12323399573676603587e-17   /* 0x3C91A62633145C07 */
#define ESM_PIO2_3  (-1.4973849048591698e-33)    /* 0xB91F1976B7ED8FBC */
#define ESM_PIO4    7.85398163397448278999e-01   /* 0x3FE921FB54442D18 */

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

/* sin(x + y) on |x + y| <= pi/4, y a tail below ulp(x) (fdlibm __kernel_sin). */
ESM_FN double esm_kernel_sin(double x, double y)
{
    const double S1 = -1.66666666666666324348e-01; /* 0xBFC5555555555549 */
    const double S2 =  8.33333333332248946124e-03; /* 0x3F8111111110F8A6 */
    const double S3 = -1.98412698298579493134e-04; /* 0xBF2A01A019C161D5 */
    const double S4 =  2.75573137070700676789e-06; /* 0x3EC71DE357B1FE7D */
    const double S5 = -2.50507602534068634195e-08; /* 0xBE5AE5E68A2B9CEB */
    const double S6 =  1.58969099521155010221e-10; /* 0x3DE5D93A5ACFD57C */
    const double z = x * x;
    const double v = z * x;
    const double r = fma(z, fma(z, fma(z, fma(z, S6, S5), S4), S3), S2);
    return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

/* cos(x + y) on |x + y| <= pi/4 (fdlibm __kernel_cos, FreeBSD form). */
ESM_FN double esm_kernel_cos(double x, double y)
{
    const double C1 =  4.16666666666666019037e-02; /* 0x3FA555555555554C */
    const double C2 = -1.38888888888741095749e-03; /* 0xBF56C16C16C15177 */
    const double C3 =  2.48015872894767294178e-05; /* 0x3EFA01A019CB1590 */
    const double C4 = -2.75573143513906633035e-07; /* 0xBE927E4F809C52AD */
    const double C5 =  2.08757232129817482790e-09; /* 0x3E21EE9EBDB4B1C4 */
    const double C6 = -1.13596475577881948265e-11; /* 0xBDA8FAE9BE8838D4 */
    const double z = x * x;
    const double r = z * fma(z, fma(z, fma(z, fma(z, fma(z, C6, C5), C4), C3), C2), C1);
    const double hz = 0.5 * z;
    const double w = 1.0 - hz;
    return w + (((1.0 - w) - hz) + (z * r - x * y));
}

/*
 * x = n*pi/2 + (hi + lo), |hi + lo| <= ~pi/4.  Returns n mod 4.
 * Step 1 (fma(-n, P1, x)) is exact: x and n*P1 are integer multiples of
 * 2^-52 for |x| >= 1, and the difference is below pi/4 in magnitude.
 */
ESM_FN int esm_rem_pio2(double x, double* hi, double* lo)
{
    if (fabs(x) <= ESM_PIO4) {
        *hi = x;
        *lo = 0.0;
        return 0;
    }
    const double n = rint(x * ESM_INVPIO2);
    const double a = fma(-n, ESM_PIO2_1, x);
    /* a - n*P2 - n*P3 in double-double. */
    const double p = n * ESM_PIO2_2;
    const double pe = fma(n, ESM_PIO2_2, -p);     /* n*P2 = p + pe exactly */
    const double s = a - p;                        /* TwoSum(a, -p) */
    const double bb = s - a;
    const double err = (a - (s - bb)) + (-p - bb);
    const double t = err - pe - n * ESM_PIO2_3;
    const double h = s + t;                        /* renormalise */
    *hi = h;
    *lo = t - (h - s);
    const int64_t ni = (int64_t)n;
    return (int)(ni & 3);
}

ESM_FN double esm_sin(double x)
{
    if (!(fabs(x) < 1073741824.0)) return x - x; /* NaN for inf/NaN/|x|>=2^30 */
    double hi, lo;
    const int q = esm_rem_pio2(x, &hi, &lo);
    switch (q) {
    case 0: return esm_kernel_sin(hi, lo);
    case 1: return esm_kernel_cos(hi, lo);
    case 2: return -esm_kernel_sin(hi, lo);
    default: return -esm_kernel_cos(hi, lo);
    }
}

ESM_FN double esm_cos(double x)
{
    if (!(fabs(x) < 1073741824.0)) return x - x;
    double hi, lo;
    const int q = esm_rem_pio2(x, &hi, &lo);
    switch (q) {
    case 0: return esm_kernel_cos(hi, lo);
    case 1: return -esm_kernel_sin(hi, lo);
    case 2: return -esm_kernel_cos(hi, lo);
    default: return esm_kernel_sin(hi, lo);
    }
}

/* sin and cos of one argument; identical to esm_sin(x), esm_cos(x). */
ESM_FN void esm_sincos(double x, double* s, double* c)
{
    if (!(fabs(x) < 1073741824.0)) {
        *s = x - x;
        *c = x - x;
        return;
    }
    double hi, lo;
    const int q = esm_rem_pio2(x, &hi, &lo);
    const double ks = esm_kernel_sin(hi, lo);
    const double kc = esm_kernel_cos(hi, lo);
    switch (q) {
    case 0: *s = ks;  *c = kc;  break;
    case 1: *s = kc;  *c = -ks; break;
    case 2: *s = -ks; *c = -kc; break;
    default: *s = -kc; *c = ks; break;
    }
}

/* asin (fdlibm e_asin.c). */
ESM_FN double esm_asin_r(double t)
{
    const double pS0 =  1.66666666666666657415e-01; /* 0x3FC5555555555555 */
    const double pS1 = -3.25565818622400915405e-01; /* 0xBFD4D61203EB6F7D */
    const double pS2 =  2.01212532134862925881e-01; /* 0x3FC9C1550E884455 */
    const double pS3 = -4.00555345006794114027e-02; /* 0xBFA48228B5688F3B */
    const double pS4 =  7.91534994289814532176e-04; /* 0x3F49EFE07501B288 */
    const double pS5 =  3.47933107596021167570e-05; /* 0x3F023DE10DFDF709 */
    const double qS1 = -2.40339491173441421878e+00; /* 0xC0033A271C8A2D4B */
    const double qS2 =  2.02094576023350569471e+00; /* 0x40002AE59C598AC8 */
    const double qS3 = -6.88283971605453293030e-01; /* 0xBFE6066C1B8D0159 */
    const double qS4 =  7.70381505559019352791e-02; /* 0x3FB3B8C5B12E9282 */
    const double p = t * fma(t, fma(t, fma(t, fma(t, fma(t, pS5, pS4), pS3), pS2), pS1), pS0);
    const double q = fma(t, fma(t, fma(t, fma(t, qS4, qS3), qS2), qS1), 1.0);
    return p / q;
}

/*
 * Branch-free arrangement of fdlibm's three ranges (|x| < 0.5, 0.5 <= |x| <
 * 0.975, 0.975 <= |x| < 1): the rational part is evaluated once, for the
 * argument of whichever range x is in, and the results are selected.  Every
 * range performs exactly the operations of e_asin.c, so the value is the
 * same; a GPU wave whose lanes fall in different ranges then runs one
 * rational evaluation instead of two.
 */
ESM_FN double esm_asin(double x)
{
    const double pio2_hi = 1.57079632679489655800e+00; /* 0x3FF921FB54442D18 */
    const double pio2_lo = 6.12323399573676603587e-17; /* 0x3C91A62633145C07 */
    const double pio4_hi = 7.85398163397448278999e-01; /* 0x3FE921FB54442D18 */
    const double ax = fabs(x);
    if (!(ax < 1.0)) {
        if (ax == 1.0) return x * pio2_hi + x * pio2_lo;
        return (x - x) / (x - x);  /* |x| > 1 or NaN: NaN */
    }
    const int small = ax < 0.5;
    const double t = small ? x * x : (1.0 - ax) * 0.5;
    const double r = esm_asin_r(t);
    /* |x| < 0.5 (below 2^-27 asin(x) rounds to x) */
    const double res_small = (ax < 7.450580596923828125e-9) ? x : fma(x, r, x);
    /* 0.5 <= |x| < 1 */
    const double s = sqrt(t);
    const double f = esm_from_bits(esm_bits(s) & 0xFFFFFFFF00000000ULL);
    const double c = (t - f * f) / (s + f);
    const double res_top = pio2_hi - (2.0 * fma(s, r, s) - pio2_lo);
    const double p = 2.0 * s * r - (pio2_lo - 2.0 * c);
    const double q = pio4_hi - 2.0 * f;
    const double res_mid = pio4_hi - (p - q);
    const double big = (ax >= 0.975) ? res_top : res_mid;
    return small ? res_small : ((x < 0.0) ? -big : big);
}

/* exp (fdlibm e_exp.c). */
ESM_FN double esm_exp(double x)
{
    const double ln2HI  = 6.93147180369123816490e-01; /* 0x3FE62E42FEE00000 */
    const double ln2LO  = 1.90821492927058770002e-10; /* 0x3DEA39EF35793C76 */
    const double invln2 = 1.44269504088896338700e+00; /* 0x3FF71547652B82FE */
    const double P1 =  1.66666666666666019037e-01;    /* 0x3FC555555555553E */
    const double P2 = -2.77777777770155933842e-03;    /* 0xBF66C16C16BEBD93 */
    const double P3 =  6.61375632143793436117e-05;    /* 0x3F11566AAF25DE2C */
    const double P4 = -1.65339022054652515390e-06;    /* 0xBEBBBD41C5D26BF1 */
    const double P5 =  4.13813679705723846039e-08;    /* 0x3E66376972BEA4D0 */
    if (x != x) return x + x;
    if (x > 7.09782712893383973096e+02) return 1.0e300 * 1.0e300; /* +inf */
    if (x < -7.45133219101941108420e+02) return 0.0;
    const double ax = fabs(x);
    if (ax < 3.7252902984e-09) return 1.0 + x;                 /* 2^-28 */
    double hi, lo, k;
    if (ax > 3.46573590279972654709e-01) {                      /* 0.5 ln2 */
        k = rint(x * invln2);
        hi = x - k * ln2HI;                                     /* exact */
        lo = k * ln2LO;
    } else {
        k = 0.0;
        hi = x;
        lo = 0.0;
    }
    const double r = hi - lo;
    const double t = r * r;
    const double c = r - t * fma(t, fma(t, fma(t, fma(t, P5, P4), P3), P2), P1);
    if (k == 0.0) return 1.0 - ((r * c) / (c - 2.0) - r);
    const double y = 1.0 - ((lo - (r * c) / (2.0 - c)) - hi);
    const int ki = (int)k;
    if (ki >= -1021) return ldexp(y, ki);
    return ldexp(y, ki + 1000) * 9.332636185032188789900895447238171696171e-302; /* 2^-1000 */
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
