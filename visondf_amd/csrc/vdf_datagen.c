/*
 * vdf_datagen.c -- seeded synthetic columns for benchmarks and tests (host,
 * OpenMP).  Same streams as visondf_amd/datagen.py: SplitMix64 keyed by
 * (seed, column id); the Python module cross-checks this library.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>

#define GOLDEN 0x9E3779B97F4A7C15ull

static inline uint64_t mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t stream_base(uint64_t seed, uint64_t col)
{
    return mix(seed * GOLDEN + (col & 0xFFFFull));
}

/* x_i = mix(s0 + (start + i + 1) * golden) */
void vdg_splitmix64(uint64_t seed, uint64_t col, uint64_t start, uint64_t n, uint64_t* out, int threads)
{
    const uint64_t s0 = stream_base(seed, col);
    #pragma omp parallel for num_threads(threads) schedule(static)
    for (uint64_t i = 0; i < n; ++i) out[i] = mix(s0 + (start + i + 1) * GOLDEN);
}

void vdg_uniform01(uint64_t seed, uint64_t col, uint64_t n, double* out, int threads)
{
    const uint64_t s0 = stream_base(seed, col);
    #pragma omp parallel for num_threads(threads) schedule(static)
    for (uint64_t i = 0; i < n; ++i) out[i] = (double)(mix(s0 + (i + 1) * GOLDEN) >> 11) * 0x1.0p-53;
}

/* Box-Muller: sqrt(-2 ln u1) cos(2 pi u2), u1 from column col, u2 from col + 1000 */
void vdg_normal01(uint64_t seed, uint64_t col, uint64_t n, double* out, int threads)
{
    const uint64_t s1 = stream_base(seed, col), s2 = stream_base(seed, col + 1000);
    const double two_pi = 2.0 * 3.141592653589793;
    #pragma omp parallel for num_threads(threads) schedule(static)
    for (uint64_t i = 0; i < n; ++i) {
        double u1 = (double)(mix(s1 + (i + 1) * GOLDEN) >> 11) * 0x1.0p-53;
        double u2 = (double)(mix(s2 + (i + 1) * GOLDEN) >> 11) * 0x1.0p-53;
        if (u1 <= 0.0) u1 = 0x1.0p-53;
        out[i] = sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
    }
}

/* Zipf(s) ranks in [0, domain): first k with cdf[k] > u (u uniform01 of the
 * stream), using a guide table over 2^22 equal-probability buckets to bound
 * the binary search. */
void vdg_zipf_ranks(uint64_t seed, uint64_t col, uint64_t n, double s, uint64_t domain, int64_t* out,
                    int threads)
{
    double* cdf = (double*)malloc(domain * sizeof(double));
    double acc = 0.0;
    for (uint64_t k = 0; k < domain; ++k) {
        acc += pow((double)(k + 1), -s);
        cdf[k] = acc;
    }
    for (uint64_t k = 0; k < domain; ++k) cdf[k] /= acc;
    const uint64_t G = 1ull << 22;
    uint64_t* guide = (uint64_t*)malloc((G + 1) * sizeof(uint64_t));
    {
        uint64_t k = 0;
        for (uint64_t b = 0; b <= G; ++b) {
            const double t = (double)b / (double)G;
            while (k < domain && !(cdf[k] > t)) ++k;
            guide[b] = k;
        }
    }
    const uint64_t s0 = stream_base(seed, col);
    #pragma omp parallel for num_threads(threads) schedule(static)
    for (uint64_t i = 0; i < n; ++i) {
        const double u = (double)(mix(s0 + (i + 1) * GOLDEN) >> 11) * 0x1.0p-53;
        uint64_t b = (uint64_t)(u * (double)G);
        if (b >= G) b = G - 1;
        uint64_t lo = guide[b], hi = guide[b + 1];  /* answer in [lo, hi] */
        if (hi > domain) hi = domain;
        while (lo < hi) {
            const uint64_t mid = lo + (hi - lo) / 2;
            if (cdf[mid] > u) hi = mid; else lo = mid + 1;
        }
        out[i] = (int64_t)(lo < domain ? lo : domain - 1);
    }
    free(guide);
    free(cdf);
}
