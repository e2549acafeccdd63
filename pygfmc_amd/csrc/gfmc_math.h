// Exactly specified FP64 functions shared by host and device code of the
// kernels (the CPU oracle restates them operation for operation).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define GFMC_MATH_HD __host__ __device__ __forceinline__
#else
#define GFMC_MATH_HD static inline
#endif

// -log(q) for q in [2^-1022, 1]: the argument reduction and the minimax
// polynomial of Sun's fdlibm e_log.c (1993): q = 2^k (1 + f) with 1 + f in
// [sqrt(2)/2, sqrt(2)), s = f / (2 + f), z = s^2, w = z^2,
// R = z (L1 + w (L3 + w (L5 + w L7))) + w (L2 + w (L4 + w L6)),
// log q = k ln2_hi - ((f^2/2 - (s (f^2/2 + R) + k ln2_lo)) - f).
// Plain IEEE operations in a fixed order (no contraction): the same bits on
// every conforming platform; error below one ulp of log q.
GFMC_MATH_HD double gfmc_neg_log_unit(double q) {
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double L1 = 6.666666666666735130e-01, L2 = 3.999999999940941908e-01;
    const double L3 = 2.857142874366239149e-01, L4 = 2.222219843214978396e-01;
    const double L5 = 1.818357216161805012e-01, L6 = 1.531383769920937332e-01;
    const double L7 = 1.479819860511658591e-01;
#if defined(__HIP_DEVICE_COMPILE__)
    const uint64_t b = (uint64_t)__double_as_longlong(q);
#else
    uint64_t b;
    __builtin_memcpy(&b, &q, sizeof b);
#endif
    const uint32_t hi = (uint32_t)(b >> 32) & 0x000fffffu;
    // 1 + f in [sqrt(2)/2, sqrt(2)): the mantissa, or half of it when it
    // lies above sqrt(2) (its high word then carries 0x3fe instead of 0x3ff)
    const uint32_t up = (hi + 0x95f64u) & 0x100000u;
    const int k = (int)((b >> 52) & 0x7ffu) - 1023 + (int)(up >> 20);
    const uint64_t mb = ((uint64_t)(hi | (up ^ 0x3ff00000u)) << 32) | (b & 0xffffffffull);
#if defined(__HIP_DEVICE_COMPILE__)
    const double m = __longlong_as_double((long long)mb);
#else
    double m;
    __builtin_memcpy(&m, &mb, sizeof m);
#endif
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    const double w = z * z;
    const double t1 = w * (L2 + w * (L4 + w * L6));
    const double t2 = z * (L1 + w * (L3 + w * (L5 + w * L7)));
    const double R = t2 + t1;
    const double hfsq = 0.5 * f * f;
    const double dk = (double)k;
    const double lg = dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
    return -lg;
}
