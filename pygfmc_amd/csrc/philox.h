// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11), host and device.
// Ten rounds of the 4x32 S-box with multipliers 0xD2511F53 / 0xCD9E8D57 and
// Weyl key increments 0x9E3779B9 / 0xBB67AE85.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define GFMC_HD __host__ __device__ __forceinline__
#else
#define GFMC_HD static inline
#endif

struct gfmc_u4 {
    uint32_t x, y, z, w;
};

GFMC_HD gfmc_u4 gfmc_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                            uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    gfmc_u4 o = {c0, c1, c2, c3};
    return o;
}

// 53-bit uniform in [0, 1) from two words (high 27 bits of a, high 26 of b).
GFMC_HD double gfmc_u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// 24-bit uniform in [0, 1).
GFMC_HD float gfmc_u24(uint32_t a) { return (float)(a >> 8) * (1.0f / 16777216.0f); }
