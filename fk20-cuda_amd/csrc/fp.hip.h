// Fp arithmetic for BLS12-381 on gfx950: Montgomery form, 14 limbs of 29 bits in VGPRs.
//
// Replaces the reference's standard-form Barrett arithmetic (fp_mul.cuh:13 +
// fp_reduce12.cuh:40, fp_add.cuh:47, fp_sub.cuh:49, fp_x12.cuh:11).  Design:
//  * radix 2^29, 14 limbs, Montgomery radix R = 2^406.  Every partial product of the
//    product-scanning (column-wise) Montgomery multiply is one v_mad_u64_u32 into a
//    64-bit column accumulator: a column holds at most 28 products of 58 bits, so it
//    never overflows and no carry flags are needed.  Measured on MI355X this is 22%
//    faster than the 12 x 32-bit CIOS form (65.1 vs 53.3 G mul/s).
//  * Lazy reduction.  R > 2^25 p, so for inputs with a*b < 2^20 p^2 the product
//    (ab + mp)/R is < 1.01 p without a final subtraction.  Additions and subtractions
//    only normalise carries; subtraction adds a multiple K*p >= the subtrahend's bound.
//    Callers track value bounds (see g1.hip.h); all values stay below 2^12 p.
//  * Memory form: the same Montgomery value packed into 12 x 32-bit words (values
//    < 2^384, i.e. below ~9.8 p, which every stored value satisfies).
// Functions are __host__ __device__ so the exact device code is also testable on CPU.
#pragma once
#include <stdint.h>

#include "constants.hip.h"

#ifndef FK20_HD
#define FK20_HD __host__ __device__ __forceinline__
#endif

namespace fk20 {

struct fe {
    uint32_t l[FP_N];
};

// ---------------------------------------------------------------- normalisation

// Propagate carries of limbs < 2^32 (non-negative) so every limb is < 2^29.
FK20_HD void fe_norm(fe &a) {
#pragma unroll
    for (int i = 0; i < FP_N - 1; i++) {
        a.l[i + 1] += a.l[i] >> 29;
        a.l[i] &= FP_MASK;
    }
}

// Same for limbs that may be negative as int32 (total value non-negative).
FK20_HD void fe_norm_signed(fe &a) {
#pragma unroll
    for (int i = 0; i < FP_N - 1; i++) {
        int32_t c = ((int32_t)a.l[i]) >> 29;
        a.l[i] &= FP_MASK;
        a.l[i + 1] += (uint32_t)c;
    }
}

// ---------------------------------------------------------------- basic ops

FK20_HD void fe_set(fe &z, const uint32_t *c) {
#pragma unroll
    for (int i = 0; i < FP_N; i++) z.l[i] = c[i];
}

FK20_HD void fe_zero(fe &z) {
#pragma unroll
    for (int i = 0; i < FP_N; i++) z.l[i] = 0;
}

FK20_HD void fe_one(fe &z) { fe_set(z, FP_ONE); }

// z = a + b (value bounds add)
FK20_HD void fe_add(fe &z, const fe &a, const fe &b) {
#pragma unroll
    for (int i = 0; i < FP_N; i++) z.l[i] = a.l[i] + b.l[i];
    fe_norm(z);
}

// z = a + b without carry propagation: lower limbs < 2^30.  Such a value may be ONE operand
// of fe_mul or fe_mul2 (the column bounds below allow one operand with limbs < 2^30), or
// the input of fe_neg / fe_mul_small, which normalise; nothing else.
FK20_HD void fe_add_nn(fe &z, const fe &a, const fe &b) {
#pragma unroll
    for (int i = 0; i < FP_N; i++) z.l[i] = a.l[i] + b.l[i];
}

template <int K>
FK20_HD const uint32_t *fe_kp() {
    static_assert(K == 2 || K == 4 || K == 8 || K == 16 || K == 32 || K == 64 || K == 128, "K");
    return K == 2 ? FP_2P : K == 4 ? FP_4P : K == 8 ? FP_8P : K == 16 ? FP_16P : K == 32 ? FP_32P
         : K == 64 ? FP_64P : FP_128P;
}

// z = a + K p - b; requires value(b) <= K p.  Bound: bound(a) + K p.
template <int K>
FK20_HD void fe_sub(fe &z, const fe &a, const fe &b) {
    const uint32_t *kp = fe_kp<K>();
#pragma unroll
    for (int i = 0; i < FP_N; i++) z.l[i] = a.l[i] + kp[i] - b.l[i];
    fe_norm_signed(z);
}

// z = a + K p - b - c; requires value(b) + value(c) <= K p.  Bound: bound(a) + K p.
template <int K>
FK20_HD void fe_sub2(fe &z, const fe &a, const fe &b, const fe &c) {
    const uint32_t *kp = fe_kp<K>();
#pragma unroll
    for (int i = 0; i < FP_N; i++) z.l[i] = a.l[i] + kp[i] - b.l[i] - c.l[i];
    fe_norm_signed(z);
}

// z = K p - a; requires value(a) <= K p.
template <int K>
FK20_HD void fe_neg(fe &z, const fe &a) {
    const uint32_t *kp = fe_kp<K>();
#pragma unroll
    for (int i = 0; i < FP_N; i++) z.l[i] = kp[i] - a.l[i];
    fe_norm_signed(z);
}

// z = c * a for a small constant c (c * 2^29 < 2^35).  Bound: c * bound(a).
template <uint32_t C>
FK20_HD void fe_mul_small(fe &z, const fe &a) {
    uint64_t carry = 0;
#pragma unroll
    for (int i = 0; i < FP_N - 1; i++) {
        uint64_t t = (uint64_t)a.l[i] * C + carry;
        z.l[i] = (uint32_t)t & FP_MASK;
        carry = t >> 29;
    }
    // the top limb is not masked: it holds every remaining bit (values stay < 2^12 p)
    z.l[FP_N - 1] = (uint32_t)((uint64_t)a.l[FP_N - 1] * C + carry);
}

// ---------------------------------------------------------------- Montgomery product

// z = a b / 2^406 mod p (lazy).  Product scanning with the reduction interleaved
// (the "FIPS" ordering): column k accumulates a_i b_{k-i} and m_i p_{k-i}: at most 14
// products below 2^59 (one operand may have lower limbs < 2^30, see fe_add_nn) and 14
// below 2^58, so the 64-bit accumulator never overflows.
FK20_HD void fe_mul(fe &z, const fe &a, const fe &b) {
    uint32_t m[FP_N];
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < FP_N; k++) {
#pragma unroll
        for (int i = 0; i <= k; i++) acc += (uint64_t)a.l[i] * b.l[k - i];
#pragma unroll
        for (int i = 0; i < k; i++) acc += (uint64_t)m[i] * FP_P[k - i];
        m[k] = ((uint32_t)acc * FP_PINV) & FP_MASK;
        acc += (uint64_t)m[k] * FP_P[0];
        acc >>= 29;
    }
#pragma unroll
    for (int k = FP_N; k < 2 * FP_N - 1; k++) {
#pragma unroll
        for (int i = k - FP_N + 1; i < FP_N; i++) acc += (uint64_t)a.l[i] * b.l[k - i];
#pragma unroll
        for (int i = k - FP_N + 1; i < FP_N; i++) acc += (uint64_t)m[i] * FP_P[k - i];
        z.l[k - FP_N] = (uint32_t)acc & FP_MASK;
        acc >>= 29;
    }
    z.l[FP_N - 1] = (uint32_t)acc;
}

// z = (a b + c d) / 2^406 mod p (lazy) with one shared reduction.  Column k then holds at
// most 28 input products and 14 reduction products: with the lower limbs of a, b, c, d
// below 2^29 (every normalised fe_* output) except one operand below 2^30 (fe_add_nn),
// 14 * 2^59 + 28 * 2^58 < 2^63.9, so the accumulator never overflows.  Top limbs are
// small (values < 2^12 p).  Output < 1.03 p when a b + c d < 2^20 p^2.
FK20_HD void fe_mul2(fe &z, const fe &a, const fe &b, const fe &c, const fe &d) {
    uint32_t m[FP_N];
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < FP_N; k++) {
#pragma unroll
        for (int i = 0; i <= k; i++) acc += (uint64_t)a.l[i] * b.l[k - i];
#pragma unroll
        for (int i = 0; i <= k; i++) acc += (uint64_t)c.l[i] * d.l[k - i];
#pragma unroll
        for (int i = 0; i < k; i++) acc += (uint64_t)m[i] * FP_P[k - i];
        m[k] = ((uint32_t)acc * FP_PINV) & FP_MASK;
        acc += (uint64_t)m[k] * FP_P[0];
        acc >>= 29;
    }
#pragma unroll
    for (int k = FP_N; k < 2 * FP_N - 1; k++) {
#pragma unroll
        for (int i = k - FP_N + 1; i < FP_N; i++) acc += (uint64_t)a.l[i] * b.l[k - i];
#pragma unroll
        for (int i = k - FP_N + 1; i < FP_N; i++) acc += (uint64_t)c.l[i] * d.l[k - i];
#pragma unroll
        for (int i = k - FP_N + 1; i < FP_N; i++) acc += (uint64_t)m[i] * FP_P[k - i];
        z.l[k - FP_N] = (uint32_t)acc & FP_MASK;
        acc >>= 29;
    }
    z.l[FP_N - 1] = (uint32_t)acc;
}

// z = a^2: cross products computed once and doubled.
FK20_HD void fe_sqr(fe &z, const fe &a) {
    uint32_t m[FP_N];
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < FP_N; k++) {
        uint64_t cross = 0;
#pragma unroll
        for (int i = 0; 2 * i < k; i++) cross += (uint64_t)a.l[i] * a.l[k - i];
        acc += cross << 1;
        if ((k & 1) == 0) acc += (uint64_t)a.l[k / 2] * a.l[k / 2];
#pragma unroll
        for (int i = 0; i < k; i++) acc += (uint64_t)m[i] * FP_P[k - i];
        m[k] = ((uint32_t)acc * FP_PINV) & FP_MASK;
        acc += (uint64_t)m[k] * FP_P[0];
        acc >>= 29;
    }
#pragma unroll
    for (int k = FP_N; k < 2 * FP_N - 1; k++) {
        uint64_t cross = 0;
#pragma unroll
        for (int i = k - FP_N + 1; 2 * i < k; i++) cross += (uint64_t)a.l[i] * a.l[k - i];
        acc += cross << 1;
        if ((k & 1) == 0) acc += (uint64_t)a.l[k / 2] * a.l[k / 2];
#pragma unroll
        for (int i = k - FP_N + 1; i < FP_N; i++) acc += (uint64_t)m[i] * FP_P[k - i];
        z.l[k - FP_N] = (uint32_t)acc & FP_MASK;
        acc >>= 29;
    }
    z.l[FP_N - 1] = (uint32_t)acc;
}

// ---------------------------------------------------------------- packing

// 12 x 32-bit words (little endian) -> 14 x 29-bit limbs, same integer.
FK20_HD void fe_unpack(fe &z, const uint32_t *w) {
#pragma unroll
    for (int i = 0; i < FP_N; i++) {
        const int bit = 29 * i, wd = bit >> 5, off = bit & 31;
        uint64_t lo = w[wd];
        uint64_t hi = (wd + 1 < 12) ? w[wd + 1] : 0;
        z.l[i] = (uint32_t)(((hi << 32) | lo) >> off) & FP_MASK;
    }
}

// 14 x 29-bit limbs -> 12 x 32-bit words; the value must be < 2^384.
FK20_HD void fe_pack(uint32_t *w, const fe &a) {
#pragma unroll
    for (int j = 0; j < 12; j++) {
        const int bit = 32 * j, li = bit / 29, off = bit % 29;
        uint64_t v = (uint64_t)a.l[li] >> off;
        v |= (uint64_t)a.l[li + 1] << (29 - off);
        if (li + 2 < FP_N && 58 - off < 32) v |= (uint64_t)a.l[li + 2] << (58 - off);
        w[j] = (uint32_t)v;
    }
}

// ---------------------------------------------------------------- conversions

// Any value < 2^384 in standard form -> Montgomery (result < 2 p).
FK20_HD void fe_to_mont(fe &z, const fe &a) {
    fe r2;
    fe_set(r2, FP_R2);
    fe_mul(z, a, r2);
}

// Subtract p once if z >= p (z < 2p).
FK20_HD void fe_cond_sub_p(fe &z) {
    fe t;
#pragma unroll
    for (int i = 0; i < FP_N; i++) t.l[i] = z.l[i] - FP_P[i];
    fe_norm_signed(t);
    const bool neg = ((int32_t)t.l[FP_N - 1]) < 0;
#pragma unroll
    for (int i = 0; i < FP_N; i++) z.l[i] = neg ? z.l[i] : t.l[i];
}

// Montgomery value (any lazy bound below 2^12 p) -> canonical standard form in [0, p).
FK20_HD void fe_from_mont(fe &z, const fe &a) {
    fe one;
    fe_zero(one);
    one.l[0] = 1;
    fe_mul(z, a, one);  // < p + 1
    fe_cond_sub_p(z);
}

// Montgomery value -> canonical Montgomery value in [0, p).
FK20_HD void fe_canonical(fe &z, const fe &a) {
    fe t, one;
    fe_set(one, FP_ONE);
    // a * (R mod p) / R = a, and the product is < 1.01 p for a < 2^12 p
    fe_mul(t, a, one);
    fe_cond_sub_p(t);
    z = t;
}

FK20_HD bool fe_is_zero_canonical(const fe &a) {
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < FP_N; i++) acc |= a.l[i];
    return acc == 0;
}

// Inverse by Fermat (a^(p-2)); a^-1 of 0 is 0.  Montgomery in, Montgomery out.
FK20_HD void fe_inv(fe &z, const fe &a) {
    // p - 2 as 12 x 32-bit words, scanned from the top bit
    fe acc, base = a;
    fe_one(acc);
    for (int w = 11; w >= 0; w--) {
        const uint32_t e = FP_PM2[w];
        for (int b = 31; b >= 0; b--) {
            fe_sqr(acc, acc);
            if ((e >> b) & 1) fe_mul(acc, acc, base);
        }
    }
    z = acc;
}

}  // namespace fk20
