// Fr arithmetic (BLS12-381 scalar field) on gfx950: Montgomery form, 8 x 32-bit limbs,
// canonical values in [0, r).
//
// Replaces the reference's standard-form Fr (fr_mul.cu:13 schoolbook + Barrett,
// fr_addsub.cu:18, fr_add.cuh, fr_sub.cuh).  Fr work on the FK20 path is the Toeplitz
// NTT only (well under 1% of the path's multiply-adds), so this is a plain CIOS with
// carry chains rather than the lazy 29-bit form used for Fp.
#pragma once
#include <stdint.h>

#include "constants.hip.h"

#ifndef FK20_HD
#define FK20_HD __host__ __device__ __forceinline__
#endif

namespace fk20 {

struct fs {
    uint32_t l[8];
};

FK20_HD uint32_t addc32(uint32_t a, uint32_t b, uint32_t cin, uint32_t &cout) {
    uint64_t s = (uint64_t)a + b + cin;
    cout = (uint32_t)(s >> 32);
    return (uint32_t)s;
}

FK20_HD uint32_t subb32(uint32_t a, uint32_t b, uint32_t bin, uint32_t &bout) {
    uint64_t d = (uint64_t)a - b - bin;
    bout = (uint32_t)(d >> 63);
    return (uint32_t)d;
}

// z = z - r if z >= r (z < 2r)
FK20_HD void fs_cond_sub(fs &z, uint32_t top) {
    fs t;
    uint32_t br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) t.l[i] = subb32(z.l[i], FR_R[i], br, br);
    // keep t unless the subtraction borrowed past the extra top word
    const bool keep = (top != 0) || (br == 0);
#pragma unroll
    for (int i = 0; i < 8; i++) z.l[i] = keep ? t.l[i] : z.l[i];
}

FK20_HD void fs_add(fs &z, const fs &a, const fs &b) {
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) z.l[i] = addc32(a.l[i], b.l[i], c, c);
    fs_cond_sub(z, c);
}

FK20_HD void fs_sub(fs &z, const fs &a, const fs &b) {
    uint32_t br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) z.l[i] = subb32(a.l[i], b.l[i], br, br);
    // add r back if negative
    uint32_t c = 0;
    const uint32_t mask = 0u - br;
#pragma unroll
    for (int i = 0; i < 8; i++) z.l[i] = addc32(z.l[i], FR_R[i] & mask, c, c);
}

// CIOS Montgomery product, inputs < r (or a < 2^256 and b < r), output < r.
FK20_HD void fs_mul(fs &z, const fs &a, const fs &b) {
    uint32_t t[10];
#pragma unroll
    for (int j = 0; j < 10; j++) t[j] = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        uint32_t c = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            uint64_t s = (uint64_t)a.l[j] * b.l[i] + t[j] + c;
            t[j] = (uint32_t)s;
            c = (uint32_t)(s >> 32);
        }
        uint64_t s = (uint64_t)t[8] + c;
        t[8] = (uint32_t)s;
        t[9] = (uint32_t)(s >> 32);
        const uint32_t m = t[0] * FR_RINV;
        s = (uint64_t)m * FR_R[0] + t[0];
        c = (uint32_t)(s >> 32);
#pragma unroll
        for (int j = 1; j < 8; j++) {
            s = (uint64_t)m * FR_R[j] + t[j] + c;
            t[j - 1] = (uint32_t)s;
            c = (uint32_t)(s >> 32);
        }
        s = (uint64_t)t[8] + c;
        t[7] = (uint32_t)s;
        t[8] = t[9] + (uint32_t)(s >> 32);
    }
#pragma unroll
    for (int j = 0; j < 8; j++) z.l[j] = t[j];
    fs_cond_sub(z, t[8]);
}

FK20_HD void fs_set(fs &z, const uint32_t *c) {
#pragma unroll
    for (int i = 0; i < 8; i++) z.l[i] = c[i];
}

FK20_HD void fs_zero(fs &z) {
#pragma unroll
    for (int i = 0; i < 8; i++) z.l[i] = 0;
}

// any 256-bit standard-form value -> its canonical representative in [0, r): 2^256 < 3 r, so
// at most two subtractions of r
FK20_HD void fs_reduce_std(fs &z) {
    fs_cond_sub(z, 0);
    fs_cond_sub(z, 0);
}

// standard form (any 256-bit value) -> canonical Montgomery
FK20_HD void fs_to_mont(fs &z, const fs &a) {
    fs r2;
    fs_set(r2, FR_R2);
    fs_mul(z, a, r2);
}

// canonical Montgomery -> canonical standard form
FK20_HD void fs_from_mont(fs &z, const fs &a) {
    fs one;
    fs_zero(one);
    one.l[0] = 1;
    fs_mul(z, a, one);
}

}  // namespace fk20
