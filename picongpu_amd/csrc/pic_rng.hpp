// Counter-based random numbers for the initial particle state.
//
// Every draw is a pure function of (seed, species, global cell, particle
// index in cell, stream), built from the SplitMix64 finaliser (Steele, Lea,
// Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014).
// Normal deviates use Marsaglia's polar method with a logarithm evaluated
// here from IEEE +,-,*,/ only (range reduction to [sqrt(1/2), sqrt(2)) and
// the series ln m = 2 atanh((m-1)/(m+1))), so host and device produce
// identical bits when compiled with -ffp-contract=off.
#pragma once

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define PIC_HD __host__ __device__
#else
#define PIC_HD
#endif

namespace pic {

PIC_HD inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

PIC_HD inline uint64_t rngKey(uint64_t seed, uint32_t species, uint64_t cell, uint32_t k, uint32_t stream) {
    uint64_t h = splitmix64(seed ^ ((uint64_t)species << 56));
    h = splitmix64(h ^ cell);
    h = splitmix64(h ^ (((uint64_t)k << 16) | (uint64_t)stream));
    return h;
}

// uniform float in [0,1) from the top 24 bits
PIC_HD inline float uniformF(uint64_t bits) { return (float)(uint32_t)(bits >> 40) * (1.0f / 16777216.0f); }

// uniform double in [0,1) from the top 53 bits
PIC_HD inline double uniformD(uint64_t bits) { return (double)(bits >> 11) * (1.0 / 9007199254740992.0); }

PIC_HD inline double detLog(double x) {
    uint64_t bits;
    memcpy(&bits, &x, sizeof(bits));
    int e = (int)((bits >> 52) & 0x7ff) - 1023;
    uint64_t const mb = (bits & 0x000FFFFFFFFFFFFFull) | (1023ull << 52);
    double m;
    memcpy(&m, &mb, sizeof(m));
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    double const s = (m - 1.0) / (m + 1.0);
    double const s2 = s * s;
    double term = s;
    double sum = s;
    for (int k = 3; k <= 33; k += 2) {
        term = term * s2;
        sum = sum + term / (double)k;
    }
    return 2.0 * sum + (double)e * 0.69314718055994530942;
}

// pair of independent standard normal deviates for one (key) stream family
PIC_HD inline void normalPair(uint64_t seed, uint32_t species, uint64_t cell, uint32_t k, uint32_t family, double* n0,
                              double* n1) {
    for (uint32_t attempt = 0; attempt < 1000u; ++attempt) {
        uint32_t const st = 8u + family * 2048u + attempt * 2u;
        double const u1 = 2.0 * uniformD(rngKey(seed, species, cell, k, st)) - 1.0;
        double const u2 = 2.0 * uniformD(rngKey(seed, species, cell, k, st + 1u)) - 1.0;
        double const s = u1 * u1 + u2 * u2;
        if (s > 0.0 && s < 1.0) {
            double const f = sqrt(-2.0 * detLog(s) / s);
            *n0 = u1 * f;
            *n1 = u2 * f;
            return;
        }
    }
    *n0 = 0.0;
    *n1 = 0.0;
}

}  // namespace pic
