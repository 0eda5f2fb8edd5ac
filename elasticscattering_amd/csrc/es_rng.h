/*
 * es_rng.h — the impurity-field generator of this implementation.
 *
 * The reference draws impurity centres with std::default_random_engine and
 * std::uniform_real_distribution (Grid.cpp:61-75), whose output is
 * implementation-defined (MSVC and libstdc++ differ), so those fields cannot
 * be reproduced.  This project defines its own counter-based generator so a
 * seed gives the same field on the CPU and on the GPU:
 *
 *   key   = mix64(seed ^ 0x6A09E667F3BCC909)
 *   z_k   = mix64(key + (k + 1) * 0x9E3779B97F4A7C15)        (k = 0, 1, ...)
 *   u_k   = (z_k >> 11) * 2^-53                                in [0, 1)
 *   x_i   = u_{2i}   * (hi - lo) + lo,   y_i = u_{2i+1} * (hi - lo) + lo
 *
 * mix64 is the SplitMix64 finaliser (Steele, Lea, Flood 2014; constants of
 * Stafford's "Mix13").  The two floating-point operations are separate
 * (callers compile with -ffp-contract=off), so x_i, y_i are bit-identical on
 * every platform.  [lo, hi] is the spawn square [-extends, size + extends].
 */
#ifndef ES_RNG_H
#define ES_RNG_H

#include <stdint.h>

#if defined(__HIPCC__)
#define ESR_FN __host__ __device__ static inline
#else
#define ESR_FN static inline
#endif

ESR_FN uint64_t esr_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

ESR_FN uint64_t esr_key(uint64_t seed)
{
    return esr_mix64(seed ^ 0x6A09E667F3BCC909ULL);
}

ESR_FN double esr_uniform(uint64_t key, uint64_t k)
{
    const uint64_t z = esr_mix64(key + (k + 1ULL) * 0x9E3779B97F4A7C15ULL);
    return (double)(z >> 11) * 1.1102230246251565404236316680908203125e-16; /* 2^-53 */
}

/* Impurity i of the field drawn from `key` in the spawn square [lo, hi]. */
ESR_FN void esr_impurity(uint64_t key, uint64_t i, double lo, double hi, double* x, double* y)
{
    const double span = hi - lo;
    const double ux = esr_uniform(key, 2ULL * i);
    const double uy = esr_uniform(key, 2ULL * i + 1ULL);
    const double sx = ux * span;
    const double sy = uy * span;
    *x = sx + lo;
    *y = sy + lo;
}

#endif /* ES_RNG_H */
