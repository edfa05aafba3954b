// Fixed-scalar G1 multiplication of the column-batched transforms: x <- k * x for a scalar
// known when the library loads (a twiddle w^e or a power 2^-j), run as a precomputed program
// over the GLV + wNAF digits of k.  Shared by the HIP kernels and, compiled by g++, by the CPU
// tests (FK20_UNIFORM is then the identity).
//
// Program encoding.  A split k lambda^-t = k1 + k2 lambda (|k1|, k2 < 2^128, phi(P) =
// (beta x, y) = lambda P, t in {0, 1, 2}): k1's digits act on phi^t(P), k2's on
// phi^(t+1)(P).  One 32-bit op per non-zero digit, read wave-uniformly:
//   bits 0..7   doublings to perform before this op (for the END op: trailing doublings)
//   bit  8      apply phi to the table entry (x times beta)
//   bit  13     apply phi^2 instead (x times beta^2)
//   bits 9..11  table entry e: the odd multiple (2e + 1) P
//   bit  12     negate
//   bit  31     END
//
// Arithmetic.  The input is converted to Jacobian coordinates and its odd multiples
// P, 3P, .., (2^(W-1) - 1) P are built with one common Z (doubling with update, then co-Z
// additions 2P + (2i - 1)P, then each entry rescaled to the last Z): on the isomorphic curve
// y^2 = x^3 + b Z^6 they are affine points, so the ladder runs Jacobian doublings (3S + 2M +
// a shared-reduction product pair) and mixed additions (4S + 5M + a pair) instead of complete
// projective formulas, and the table holds x and y only.  At the end the accumulator's Z is
// multiplied by the table's Z and the point converted back to homogeneous coordinates.
//
// The formulas are incomplete.  For an input of prime order r that is not infinity, the
// accumulator before each addition is a P and the entry b P with a, b known scalars; the host
// checks every program (ctx.cpp) for a != 0 and a != +-b mod r, so no exceptional case
// occurs.  An input at infinity (Z = 0) is detected and returned as infinity.  Inputs outside
// the order-r subgroup are not handled here (the GLV split assumes phi(P) = lambda P): the
// transforms find such rows with g1_in_subgroup below and recompute them on another path.
#pragma once
#include "g1.hip.h"

#ifndef FK20_UNIFORM
#define FK20_UNIFORM(x) __builtin_amdgcn_readfirstlane(x)
#endif

namespace fk20 {

static constexpr uint32_t TWOP_DBL_MASK = 0xffu;
static constexpr uint32_t TWOP_PHI = 1u << 8;
static constexpr unsigned TWOP_ENTRY_SHIFT = 9;
static constexpr uint32_t TWOP_NEG = 1u << 12;
static constexpr uint32_t TWOP_PHI2 = 1u << 13;
static constexpr uint32_t TWOP_END = 1u << 31;

// wNAF width used by the G1 transform kernels (odd multiples P, 3P, .., 15P, kept in private
// memory); the twiddle programs are built for the same width
#ifndef FK20_TW_WNAF
#define FK20_TW_WNAF 5
#endif
static constexpr int TW_WNAF = FK20_TW_WNAF;

// x <- k * x for the wave-uniform scalar whose program starts at ops[0].  x: homogeneous
// projective, coordinates < 10 p; out: x < 2 p, y < 4 p, z < 2 p.
template <int W>
FK20_HD void tw_mul(g1p &x, const uint32_t *ops) {
    constexpr int NT = 1 << (W - 2);
    static_assert(NT >= 2, "table of at least P and 3P");
    // The table lives in the lane's private memory, not in registers: each addition reads
    // one entry chosen by a wave-uniform index, so the loads are coalesced per dword and the
    // ladder's register footprint leaves room for several waves per SIMD.
    fe TX[NT], TY[NT], MU[NT - 1];
    fe zc;
    fe_canonical(zc, x.z);
    const bool inf = fe_is_zero_canonical(zc);
    // homogeneous (X : Y : Z) -> Jacobian (X Z, Y Z^2, Z)
    g1j P;
    fe_sqr(zc, x.z);                // < 2p
    fe_mul(P.x, x.x, x.z);          // < 2p
    fe_mul(P.y, x.y, zc);           // < 2p
    P.z = x.z;                      // < 10p
    g1j Q;                          // 2P, re-scaled by every co-Z addition
    g1j_dblu(Q, TX[0], TY[0], P);
    fe x1 = Q.x, y1 = Q.y;
#pragma unroll
    for (int k = 1; k < NT; k++) g1j_zaddu(TX[k], TY[k], MU[k - 1], x1, y1, TX[k - 1], TY[k - 1]);
    // entry k has Z0 * mu_0 ... mu_{k-1}; bring it to the last entry's Z by the product of
    // the later factors
    fe s = MU[NT - 2], s2, s3;
#pragma unroll
    for (int k = NT - 2; k >= 0; k--) {
        if (k < NT - 2) fe_mul(s, s, MU[k]);   // < 2p
        fe_sqr(s2, s);
        fe_mul(s3, s2, s);
        fe_mul(TX[k], TX[k], s2);
        fe_mul(TY[k], TY[k], s3);
    }
    fe zt;
    fe_mul(zt, Q.z, s);             // the table's common Z < 2p
    fe beta, beta2;
    fe_set(beta, FP_BETA);
    fe_set(beta2, FP_BETA2);
    g1j acc;
    bool started = false;
#pragma unroll 1
    for (unsigned k = 0;; k++) {
        const uint32_t op = FK20_UNIFORM(ops[k]);
        const unsigned nd = op & TWOP_DBL_MASK;
#pragma unroll 1
        for (unsigned d = 0; d < nd; d++) g1j_dbl(acc, acc);
        if (op & TWOP_END) break;
        const unsigned e = (op >> TWOP_ENTRY_SHIFT) & (NT - 1);
        fe qx = TX[e], qy = TY[e];
        if (op & TWOP_PHI2) fe_mul(qx, qx, beta2);
        else if (op & TWOP_PHI) fe_mul(qx, qx, beta);
        if (op & TWOP_NEG) fe_neg<4>(qy, qy);
        if (started) {
            g1j_madd(acc, acc, qx, qy);
        } else {
            acc.x = qx;
            acc.y = qy;
            fe_one(acc.z);
            started = true;
        }
    }
    // Jacobian with Z = acc.z * zt -> homogeneous (X Z : Y : Z^3)
    fe z, z2;
    fe_mul(z, acc.z, zt);
    fe_sqr(z2, z);
    fe_mul(x.z, z2, z);
    fe_mul(x.x, acc.x, z);
    x.y = acc.y;
    if (inf) g1_inf(x);
}

// x <- [w^384] x = [w^-128] x (w128 = false) or [w^128] x (w128 = true) for x in G1.  With
// z = -0xd201000000010000 the curve parameter and lambda = z^2 - 1 the eigenvalue of phi,
// w^128 = -z^3 = |z| (lambda + 1) mod r, and 1 + lambda + lambda^2 = 0 on G1 gives
// [w^128] x = -[|z|] phi^2(x) = -phi^2([|z|] x), [w^-128] x = phi^2([|z|] x).  [|z|] x is 63
// Jacobian doublings and 5 mixed additions of x (the set bits 62, 60, 57, 48, 16 of |z| below
// the top one), x affine on the isomorphic curve of its own Z as in tw_mul; phi^2 multiplies
// the result's X by beta^2.  This replaces a program of two 64-bit halves, its 8-entry table
// and about 11 additions.  No exceptional case for x in G1 \ {O}: before each addition the
// accumulator is [m] x with m even, 2 <= m < 2^64, so m -+ 1 is not a multiple of r.
// x: homogeneous projective, coordinates < 10 p; out: x < 2 p, y < 4 p, z < 2 p.
FK20_HD void tw_mul_z3(g1p &x, bool w128) {
    constexpr uint64_t ZABS = 0xd201000000010000ull;
    fe zc;
    fe_canonical(zc, x.z);
    const bool inf = fe_is_zero_canonical(zc);
    fe px, py;
    fe_sqr(zc, x.z);                // < 2p
    fe_mul(px, x.x, x.z);           // Jacobian X = X Z < 2p
    fe_mul(py, x.y, zc);            // Jacobian Y = Y Z^2 < 2p
    g1j acc;
    acc.x = px;
    acc.y = py;
    fe_one(acc.z);
#pragma unroll 1
    for (int b = 62; b >= 0; b--) {
        g1j_dbl(acc, acc);
        if ((ZABS >> b) & 1) g1j_madd(acc, acc, px, py);
    }
    // Jacobian with Z = acc.z * x.z -> homogeneous (X Z beta^2 : +-Y : Z^3)
    fe z, z2, beta2;
    fe_mul(z, acc.z, x.z);
    fe_sqr(z2, z);
    fe_mul(x.z, z2, z);
    fe_set(beta2, FP_BETA2);
    fe_mul(z, z, beta2);
    fe_mul(x.x, acc.x, z);
    if (w128) fe_neg<4>(x.y, acc.y);
    else x.y = acc.y;
    if (inf) g1_inf(x);
}

// A point held in registers (the host build; the kernel keeps its copies in LDS instead, to
// leave the registers to the formulas).
struct RegSlot {
    g1p v;
    FK20_HD void put(const g1p &p) { v = p; }
    FK20_HD void get(g1p &p) const { p = v; }
};

// Whether p lies in the order-r subgroup G1 (infinity included): phi^2(p) = (beta^2 x, y)
// acts on G1 as lambda^2 = -z^2, and a point of E(Fp) satisfies phi^2(p) + [z^2] p = 0 only
// if it lies in G1 (Scott, "A note on group membership tests for G1, G2 and GT on BLS
// pairing-friendly curves", 2021).  [z^2] p = [|z|]([|z|] p) by double-and-add over
// |z| = 0xd201000000010000 (63 doublings and 5 additions per round), complete formulas.
// `orig` and `base` hold p and each round's base while the doublings run.
template <class Slot>
FK20_HD bool g1_in_subgroup(const g1p &p, Slot &orig, Slot &base) {
    constexpr uint64_t ZABS = 0xd201000000010000ull;
    orig.put(p);
    g1p q = p;
#pragma unroll 1
    for (int round = 0; round < 2; round++) {
        base.put(q);
#pragma unroll 1
        for (int b = 62; b >= 0; b--) {
            g1_dbl(q, q);
            if ((ZABS >> b) & 1) {
                g1p t;
                base.get(t);
                g1_add(q, q, t);
            }
        }
    }
    g1p f;
    orig.get(f);
    fe beta2;
    fe_set(beta2, FP_BETA2);
    fe_mul(f.x, f.x, beta2);
    g1_add(q, q, f);
    fe zc;
    fe_canonical(zc, q.z);
    return fe_is_zero_canonical(zc);
}

FK20_HD bool g1_in_subgroup(const g1p &p) {
    RegSlot orig, base;
    return g1_in_subgroup(p, orig, base);
}

}  // namespace fk20
