// Elementary functions used by the FP64 flux path, written so that the host
// (g++) and gfx950 (hipcc) builds return bit-identical results: they use only
// IEEE-754 correctly rounded +, -, *, /, sqrt, exact power-of-two scaling and
// integer bit manipulation, and every translation unit that includes this
// header is compiled with -ffp-contract=off.
//
//   afar_exp   Cody-Waite argument reduction (W. J. Cody and W. Waite,
//              "Software Manual for the Elementary Functions", 1980) with the
//              fdlibm split of ln 2, then the degree-13 Taylor polynomial of
//              exp(r), |r| <= ln2/2 (truncation error < 4e-18), and exact
//              scaling by 2^n.
//   afar_erfc  Schonfelder's Chebyshev expansion of (1+2x) exp(x^2) erfc(x)
//              in t = (x-K)/(x+K) (coefficients in afar_erfc_table.h), times
//              exp(-x^2) evaluated with the high/low split of x so that the
//              rounding of x^2 does not enter; erfc(-x) = 2 - erfc(x).
//   afar_hypot sqrt(x*x + y*y); the reference calls CUDA's hypot
//              (src/CUDA/cudaEuler2DmixO3RK3/cudaFieldOperation.h:145-146,
//              cudaLimiter.h:261) -- one definition is shared by every caller.
//   afar_expf, afar_erfcf
//              the same two methods in IEEE single precision for the FP32-mixed
//              flux: the float split of ln 2 (ln2_hi with 15 significant bits,
//              so n * ln2_hi is exact for |n| < 512), the degree-7 Taylor
//              polynomial (truncation error < 6e-9 on |r| <= ln2/2), and
//              Schonfelder's series to degree 13 in float (the terms dropped
//              sum to < 1e-9); exp(-x^2) with a 12-bit high part of x.
#pragma once
#include <stdint.h>

#include "afar_erfc_table.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define AFAR_HD __host__ __device__ inline
#else
#include <math.h>
#define AFAR_HD static inline
#endif

AFAR_HD uint64_t afar_bits(double x) { uint64_t u; __builtin_memcpy(&u, &x, 8); return u; }
AFAR_HD double afar_from_bits(uint64_t u) { double x; __builtin_memcpy(&x, &u, 8); return x; }

// 2^n for -1022 <= n <= 1023, built from its exponent field (exact).
AFAR_HD double afar_pow2i(int n) { return afar_from_bits((uint64_t)(n + 1023) << 52); }

AFAR_HD double afar_hypot(double x, double y) { return sqrt(x * x + y * y); }

AFAR_HD double afar_exp(double x) {
    if (x != x) return x;                       // NaN
    if (x > 709.782712893384) return afar_from_bits(0x7ff0000000000000ULL);
    if (x < -745.1332191019412) return 0.0;
    const double log2e = 1.4426950408889634;
    const double ln2_hi = 6.93147180369123816490e-01;  // 0x3fe62e42fee00000
    const double ln2_lo = 1.90821492927058770002e-10;  // 0x3dea39ef35793c76
    // n = nearest integer to x*log2(e); |n| <= 1075 so n*ln2_hi is exact.
    double t = x * log2e;
    double n = (t >= 0.0) ? (double)(int64_t)(t + 0.5) : -(double)(int64_t)(0.5 - t);
    const double r = (x - n * ln2_hi) - n * ln2_lo;
    // Taylor polynomial of exp(r) to degree 13, Horner form.
    double p = 1.0 / 6227020800.0;          // 1/13!
    p = p * r + 1.0 / 479001600.0;          // 1/12!
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    int ni = (int)n;
    if (ni > 1023) {                 // exp(x) close to overflow: split the scaling
        p = p * 2.0;
        ni -= 1;
    }
    if (ni < -1000) {                // result may be subnormal: one rounding, at the end
        p = p * afar_pow2i(-1000 + 0);  // exact (p >= 0.7, stays normal)
        ni += 1000;
        return p * afar_pow2i(ni);
    }
    return p * afar_pow2i(ni);
}

// exp(-x*x) for x >= 0 without the rounding error of x*x.
AFAR_HD double afar_exp_negsq(double x) {
    const double xh = afar_from_bits(afar_bits(x) & 0xfffffffff8000000ULL);  // 26 bits
    const double d = (x - xh) * (x + xh);                                    // x^2 - xh^2
    return afar_exp(-(xh * xh)) * afar_exp(-d);
}

AFAR_HD double afar_erfc(double x) {
    if (x != x) return x;
    const double ax = (x < 0.0) ? -x : x;
    double r;
    if (ax > 27.5) {
        r = 0.0;                      // erfc(27.5) < 2^-1074
    } else {
        const double t = (ax - AFAR_ERFCX_K) / (ax + AFAR_ERFCX_K);
        double g;
        AFAR_ERFCX_CLENSHAW(t, g);
        r = afar_exp_negsq(ax) * (g / (1.0 + 2.0 * ax));
    }
    return (x < 0.0) ? 2.0 - r : r;
}

// ---------------------------------------------------------------- single precision

AFAR_HD uint32_t afar_bitsf(float x) { uint32_t u; __builtin_memcpy(&u, &x, 4); return u; }
AFAR_HD float afar_from_bitsf(uint32_t u) { float x; __builtin_memcpy(&x, &u, 4); return x; }

// 2^n for -126 <= n <= 127 (exact).
AFAR_HD float afar_pow2if(int n) { return afar_from_bitsf((uint32_t)(n + 127) << 23); }

AFAR_HD float afar_expf(float x) {
    if (x != x) return x;                          // NaN
    if (x > 89.0f) return afar_from_bitsf(0x7f800000u);
    if (x < -104.0f) return 0.0f;                  // exp(-104) < half the smallest subnormal
    const float log2e = 1.44269504f;
    const float ln2_hi = 0.693145751953125f;       // 0x3f317200: 15 significant bits
    const float ln2_lo = 1.42860677e-06f;          // ln 2 - ln2_hi
    const float t = x * log2e;
    const float n = (t >= 0.0f) ? (float)(int)(t + 0.5f) : -(float)(int)(0.5f - t);
    const float r = (x - n * ln2_hi) - n * ln2_lo;
    float p = 1.0f / 5040.0f;                      // 1/7!
    p = p * r + 1.0f / 720.0f;
    p = p * r + 1.0f / 120.0f;
    p = p * r + 1.0f / 24.0f;
    p = p * r + 1.0f / 6.0f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    int ni = (int)n;
    if (ni > 127) {                                // close to overflow: split the scaling
        p = p * 2.0f;
        ni -= 1;
    }
    if (ni < -125) {                               // result may be subnormal: one rounding, at the end
        p = p * afar_pow2if(-100);                 // exact (p >= 0.7 stays normal)
        ni += 100;
    }
    return p * afar_pow2if(ni);
}

// exp(-x*x) for x >= 0 without the rounding error of x*x.
AFAR_HD float afar_exp_negsqf(float x) {
    const float xh = afar_from_bitsf(afar_bitsf(x) & 0xfffff000u);  // 12 significant bits: xh*xh exact
    const float d = (x - xh) * (x + xh);                            // x^2 - xh^2
    return afar_expf(-(xh * xh)) * afar_expf(-d);
}

AFAR_HD float afar_erfcf(float x) {
    if (x != x) return x;
    const float ax = (x < 0.0f) ? -x : x;
    float r;
    if (ax > 10.1f) {
        r = 0.0f;                                  // erfc(10.1) < half the smallest subnormal
    } else {
        const float t = (ax - (float)AFAR_ERFCX_K) / (ax + (float)AFAR_ERFCX_K);
        float g;
        AFAR_ERFCX_CLENSHAW_F(t, g);
        r = afar_exp_negsqf(ax) * (g / (1.0f + 2.0f * ax));
    }
    return (x < 0.0f) ? 2.0f - r : r;
}
