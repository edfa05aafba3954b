// G1 arithmetic of BLS12-381 (y^2 = x^3 + 4) in homogeneous projective coordinates.
//
// Complete formulas of Renes, Costello and Batina (eprint 2015/1060): algorithm 7
// (addition), 8 (mixed addition with an affine point) and 9 (doubling) for a = 0,
// b3 = 3b = 12 -- the same formulas the reference uses (g1p_add.cu:78, g1p_dbl.cu:28,
// g1p_multi.cu:83), written for the lazy 29-bit Fp of fp.hip.h.  The point at infinity is
// (0 : 1 : 0); any (0 : y != 0 : 0) is accepted as infinity (g1.cuh:14 G1P_ANYINF).
//
// Value bounds (multiples of p) are annotated per step.  Inputs may be anything
// below 10 p (every packed value is < 2^384 < 9.9 p; negated y < 16 p and additions
// accept up to 20 p); outputs are below 4 p.  The
// subtrahend bound decides each fe_sub<K> constant; every product has a*b < 2^20 p^2.
#pragma once
#include "fp.hip.h"
#include "fr.hip.h"

namespace fk20 {

struct g1p {
    fe x, y, z;
};

// affine point (x, y); infinity is not representable and is handled by callers
struct g1a {
    fe x, y;
};

FK20_HD void g1_inf(g1p &p) {
    fe_zero(p.x);
    fe_one(p.y);
    fe_zero(p.z);
}

// RCB algorithm 7: r = p + q.  Inputs < 20 p, output < 2 p.  The three closing
// sums of products (X3, Y3, Z3) each share one Montgomery reduction (fe_mul2).
FK20_HD void g1_add(g1p &r, const g1p &p, const g1p &q) {
    fe t0, t1, t2, t3, t4, X3, Y3, Z3;
    fe_mul(t0, p.x, q.x);           // t0 < 2p
    fe_mul(t1, p.y, q.y);           // t1 < 2p
    fe_mul(t2, p.z, q.z);           // t2 < 2p
    fe_add_nn(t3, p.x, p.y);        // < 40p, limbs < 2^30
    fe_add(t4, q.x, q.y);           // < 40p
    fe_mul(t3, t3, t4);             // 1600p^2 -> < 2p
    fe_sub2<4>(t3, t3, t0, t1);     // X1 Y2 + X2 Y1 < 6p
    fe_add_nn(t4, p.y, p.z);        // < 40p, limbs < 2^30
    fe_add(X3, q.y, q.z);           // < 40p
    fe_mul(t4, t4, X3);             // < 2p
    fe_sub2<4>(t4, t4, t1, t2);     // Y1 Z2 + Y2 Z1 < 6p
    fe_add_nn(X3, p.x, p.z);        // < 40p, limbs < 2^30
    fe_add(Y3, q.x, q.z);           // < 40p
    fe_mul(X3, X3, Y3);             // < 2p
    fe_sub2<4>(Y3, X3, t0, t2);     // X1 Z2 + X2 Z1 < 6p
    fe_mul_small<3>(t0, t0);        // < 6p
    fe_mul_small<12>(t2, t2);       // < 24p
    fe_add_nn(Z3, t1, t2);          // < 26p, limbs < 2^30
    fe_sub<32>(t1, t1, t2);         // < 34p
    fe_mul_small<12>(Y3, Y3);       // < 72p
    fe_neg<8>(X3, t4);              // -t4 < 8p
    fe_mul2(X3, t3, t1, X3, Y3);    // t3 t1 - t4 Y3: 6*34 + 8*72 = 780p^2 -> < 2p
    fe_mul2(Y3, Y3, t0, t1, Z3);    // 72*6 + 34*26 = 1316p^2 -> < 2p (Z3 the one nn operand)
    fe_mul2(Z3, Z3, t4, t0, t3);    // 26*6 + 6*6 = 192p^2 -> < 2p (Z3 the one nn operand)
    r.x = X3;
    r.y = Y3;
    r.z = Z3;
}

// RCB algorithm 8: r = p + (x2, y2), affine operand not infinity.  Inputs < 10 p,
// output < 2 p.
FK20_HD void g1_add_aff(g1p &r, const g1p &p, const fe &X2, const fe &Y2) {
    fe t0, t1, t2, t3, t4, X3, Y3, Z3;
    fe_mul(t0, p.x, X2);            // < 2p
    fe_mul(t1, p.y, Y2);            // < 2p
    fe_add_nn(t3, X2, Y2);          // < 20p, limbs < 2^30
    fe_add(t4, p.x, p.y);           // < 20p
    fe_mul(t3, t3, t4);             // < 2p
    fe_sub2<4>(t3, t3, t0, t1);     // < 6p
    fe_mul(t4, Y2, p.z);            // < 2p
    fe_add_nn(t4, t4, p.y);         // < 12p, limbs < 2^30
    fe_mul(Y3, X2, p.z);            // < 2p
    fe_add_nn(Y3, Y3, p.x);         // < 12p, limbs < 2^30
    fe_mul_small<3>(t0, t0);        // < 6p
    fe_mul_small<12>(t2, p.z);      // < 120p
    fe_add(Z3, t1, t2);             // < 122p
    fe_sub<128>(t1, t1, t2);        // < 130p
    fe_mul_small<12>(Y3, Y3);       // < 144p (normalised)
    fe_neg<16>(X3, t4);             // -t4 < 16p (normalised)
    fe_mul2(X3, t3, t1, X3, Y3);    // 6*130 + 16*144 = 3084p^2 -> < 2p
    fe_mul2(Y3, Y3, t0, t1, Z3);    // 144*6 + 130*122 = 16724p^2 -> < 2p
    fe_mul2(Z3, Z3, t4, t0, t3);    // 122*12 + 6*6 = 1500p^2 -> < 2p (t4 the one nn operand)
    r.x = X3;
    r.y = Y3;
    r.z = Z3;
}

// RCB algorithm 9: r = 2p.  Input < 20 p, output X < 4 p, Y, Z < 2 p.
FK20_HD void g1_dbl(g1p &r, const g1p &p) {
    fe t0, t1, t2, X3, Y3, Z3;
    fe_sqr(t0, p.y);                // Y^2 < 2p
    fe_mul_small<8>(Z3, t0);        // 8 Y^2 < 16p
    fe_mul(t1, p.y, p.z);           // < 2p
    fe_sqr(X3, p.z);                // Z^2 < 2p
    fe_mul_small<12>(t2, X3);       // 3b Z^2 < 24p
    fe_mul_small<36>(X3, X3);       // 3 t2 < 72p
    fe_add_nn(Y3, t0, t2);          // < 26p, limbs < 2^30
    fe_sub<128>(t0, t0, X3);        // Y^2 - 3 t2 < 130p
    fe_mul2(Y3, t2, Z3, t0, Y3);    // t2 Z3 + t0 Y3: 24*16 + 130*26 = 3764p^2 -> < 2p
    fe_mul(Z3, t1, Z3);             // < 2p
    fe_mul(t1, p.x, p.y);           // < 2p
    fe_mul(X3, t0, t1);             // < 2p
    fe_add(X3, X3, X3);             // < 4p
    r.x = X3;
    r.y = Y3;
    r.z = Z3;
}

// r = -p.  Input < 10 p (y bound), output y < 16 p... kept below 10p by using 16p - y
// only where y < 16p; callers pass values < 4p or < 10p.
FK20_HD void g1_neg(g1p &r, const g1p &p) {
    r.x = p.x;
    fe_neg<16>(r.y, p.y);
    r.z = p.z;
}

// (p, q) <- (p + q, p - q)
FK20_HD void g1_addsub(g1p &p, g1p &q) {
    g1p nq;
    g1_neg(nq, q);      // y < 16p: g1_add accepts inputs < 20p in practice (see bound notes)
    g1p s;
    g1_add(s, p, q);
    g1_add(q, p, nq);
    p = s;
}

// ---------------------------------------------------------------- memory forms

// Standard-form g1p_t (reference ABI: 3 x 6 x u64, little endian) <-> Montgomery registers.
FK20_HD void g1_load_std(g1p &p, const uint32_t *w /* 36 words */) {
    fe t;
    fe_unpack(t, w);
    fe_to_mont(p.x, t);
    fe_unpack(t, w + 12);
    fe_to_mont(p.y, t);
    fe_unpack(t, w + 24);
    fe_to_mont(p.z, t);
}

// Canonical standard form (each coordinate < p).
FK20_HD void g1_store_std(uint32_t *w, const g1p &p) {
    fe t;
    fe_from_mont(t, p.x);
    fe_pack(w, t);
    fe_from_mont(t, p.y);
    fe_pack(w + 12, t);
    fe_from_mont(t, p.z);
    fe_pack(w + 24, t);
}

// Packed Montgomery form (internal buffers): 36 words, values < 2^384.
FK20_HD void g1_load_packed(g1p &p, const uint32_t *w) {
    fe_unpack(p.x, w);
    fe_unpack(p.y, w + 12);
    fe_unpack(p.z, w + 24);
}

FK20_HD void g1_store_packed(uint32_t *w, const g1p &p) {
    fe_pack(w, p.x);
    fe_pack(w + 12, p.y);
    fe_pack(w + 24, p.z);
}

// ---------------------------------------------------------------- scalar multiplication

// r = k p by MSB-first double-and-add over the 256-bit standard-form scalar k
// (8 x 32-bit words), as g1p_mul.cu:18-66.  k need not be reduced mod r.
FK20_HD void g1_mul(g1p &r, const g1p &p, const uint32_t *k) {
    g1p q;
    g1_inf(q);
    for (int w = 7; w >= 0; w--) {
        const uint32_t kw = k[w];
        for (int b = 31; b >= 0; b--) {
            g1_dbl(q, q);
            if ((kw >> b) & 1) g1_add(q, q, p);
        }
    }
    r = q;
}

// Affine conversion: (x/z, y/z) canonical Montgomery, or the infinity flag when z = 0.
FK20_HD bool g1_to_affine(fe &ax, fe &ay, const g1p &p) {
    fe zc;
    fe_canonical(zc, p.z);
    if (fe_is_zero_canonical(zc)) {
        fe_zero(ax);
        fe_zero(ay);
        return true;
    }
    fe zi;
    fe_inv(zi, zc);
    fe_mul(ax, p.x, zi);
    fe_mul(ay, p.y, zi);
    fe_canonical(ax, ax);
    fe_canonical(ay, ay);
    return false;
}

// 48-byte compressed serialisation, as 12 words in memory order: big-endian canonical x, top
// byte flags 0x80 (compressed), 0x40 (infinity: 0xc0 then zeros), 0x20 (y is the larger of
// y and p - y).
FK20_HD void g1_compress(uint32_t *w, const g1p &p) {
    fe x, y;
    if (g1_to_affine(x, y, p)) {
#pragma unroll
        for (int i = 0; i < 12; i++) w[i] = 0;
        w[0] = 0xc0u;
        return;
    }
    fe_from_mont(x, x);
    fe_from_mont(y, y);
    uint32_t xw[12], yw[12];
    fe_pack(xw, x);
    fe_pack(yw, y);
    // y > (p - 1) / 2, compared from the top word
    int cmp = 0;
#pragma unroll
    for (int i = 11; i >= 0; i--)
        if (cmp == 0) cmp = yw[i] > FP_HALF_PACKED[i] ? 1 : yw[i] < FP_HALF_PACKED[i] ? -1 : 0;
#pragma unroll
    for (int k = 0; k < 12; k++) w[k] = __builtin_bswap32(xw[11 - k]);
    w[0] |= 0x80u | (cmp > 0 ? 0x20u : 0u);
}

}  // namespace fk20
