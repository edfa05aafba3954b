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

// ---------------------------------------------------------------- Jacobian (ladders)
//
// Jacobian coordinates (x = X/Z^2, y = Y/Z^3) for the fixed-scalar ladders of the G1
// transforms.  These formulas are incomplete: callers guarantee that no operand is infinity
// and that no addition sees equal or opposite points (the twiddle programs are checked for
// that on the host; see ladder.hip.h).  None of them uses the curve constant b, so they also
// run on the isomorphic curve y^2 = x^3 + b u^6 on which a co-Z table's points are affine.
struct g1j {
    fe x, y, z;
};

// dbl-2009-l (Bernstein-Lange EFD, a = 0) with D = 4 X Y^2 from one product and
// Y3 = E (D - X3) - 8 Y^4 as one shared-reduction sum of products: 3S + 2M + 1 fe_mul2.
// In: X < 32 p, Y < 4 p, Z < 8 p.  Out: X < 18 p, Y < 2 p, Z < 4 p.
FK20_HD void g1j_dbl(g1j &r, const g1j &p) {
    fe A, B, D, E, F, t, u, Z3;
    fe_sqr(A, p.x);                 // X^2 < 2p
    fe_sqr(B, p.y);                 // Y^2 < 2p
    fe_mul(D, p.x, B);              // < 2p
    fe_mul_small<4>(D, D);          // D = 4 X Y^2 < 8p
    fe_mul(Z3, p.y, p.z);           // 4*8 = 32p^2 -> < 2p
    fe_add(Z3, Z3, Z3);             // 2 Y Z < 4p
    fe_mul_small<3>(E, A);          // E = 3 X^2 < 6p
    fe_sqr(F, E);                   // < 2p
    fe_add(t, D, D);                // 2D < 16p
    fe_sub<16>(r.x, F, t);          // X3 = F - 2D < 18p
    fe_sub<32>(t, D, r.x);          // D - X3 < 40p
    fe_mul_small<8>(u, B);          // 8 Y^2 < 16p
    fe_neg<16>(u, u);               // -8 Y^2 < 16p
    fe_mul2(r.y, E, t, B, u);       // E (D - X3) - 8 Y^4: 6*40 + 2*16 = 272p^2 -> < 2p
    r.z = Z3;
}

// madd-2007-bl: r = p + (x2, y2), the affine operand on the same (possibly scaled) curve,
// neither operand infinity, (x2, y2) != +-p.  4S + 5M + 1 fe_mul2.
// In: X < 32 p, Y < 4 p, Z < 8 p; x2 < 6 p, y2 < 4 p.  Out: X < 10 p, Y < 2 p, Z < 6 p.
FK20_HD void g1j_madd(g1j &r, const g1j &p, const fe &x2, const fe &y2) {
    fe Z1Z1, U2, S2, H, HH, I, J, R, V, t;
    fe_sqr(Z1Z1, p.z);              // < 2p
    fe_mul(U2, x2, Z1Z1);           // < 2p
    fe_mul(S2, p.z, Z1Z1);          // Z^3 < 2p
    fe_mul(S2, y2, S2);             // < 2p
    fe_sub<32>(H, U2, p.x);         // H = U2 - X1 < 34p
    fe_sqr(HH, H);                  // 34^2 = 1156p^2 -> < 2p
    fe_mul_small<4>(I, HH);         // I = 4 H^2 < 8p
    fe_mul(J, H, I);                // 272p^2 -> < 2p
    fe_sub<4>(R, S2, p.y);          // < 6p
    fe_add(R, R, R);                // R = 2 (S2 - Y1) < 12p
    fe_mul(V, p.x, I);              // 256p^2 -> < 2p
    fe_add(t, V, V);                // 2V < 4p
    fe_add(r.z, p.z, H);            // Z1 + H < 42p
    fe_sqr(r.z, r.z);               // 1764p^2 -> < 2p
    fe_sub2<4>(r.z, r.z, Z1Z1, HH); // Z3 = (Z1 + H)^2 - Z1Z1 - HH < 6p
    fe sy;
    fe_sqr(sy, R);                  // 144p^2 -> < 2p
    fe_sub2<8>(sy, sy, J, t);       // X3 = R^2 - J - 2V < 10p
    fe_sub<16>(V, V, sy);           // V - X3 < 18p
    fe_add(J, J, J);                // 2J < 4p
    fe_neg<4>(J, J);                // -2J < 4p
    fe_mul2(r.y, R, V, p.y, J);     // R (V - X3) - 2 Y1 J: 12*18 + 4*4 = 232p^2 -> < 2p
    r.x = sy;
}

// Doubling with update (the co-Z table start): q = 2p and (x1, y1) = p, both with
// Z = 2 Y Z.  4S + 3M.  In: X < 32 p, Y < 4 p, Z < 10 p.  Out: q.x, q.y < 18 p, q.z < 4 p;
// x1 < 8 p, y1 < 16 p.
FK20_HD void g1j_dblu(g1j &q, fe &x1, fe &y1, const g1j &p) {
    fe A, B, C, D, E, F, t;
    fe_sqr(A, p.x);                 // < 2p
    fe_sqr(B, p.y);                 // < 2p
    fe_sqr(C, B);                   // Y^4 < 2p
    fe_mul(D, p.x, B);              // < 2p
    fe_mul_small<4>(D, D);          // D = 4 X Y^2 = X (2Y)^2 < 8p
    fe_mul(q.z, p.y, p.z);          // 40p^2 -> < 2p
    fe_add(q.z, q.z, q.z);          // 2 Y Z < 4p
    fe_mul_small<3>(E, A);          // < 6p
    fe_sqr(F, E);                   // < 2p
    fe_add(t, D, D);                // < 16p
    fe_sub<16>(q.x, F, t);          // X3 < 18p
    fe_sub<32>(t, D, q.x);          // < 40p
    fe_mul(q.y, E, t);              // < 2p
    fe_mul_small<8>(C, C);          // 8 Y^4 = Y (2Y)^3 < 16p
    fe_sub<16>(q.y, q.y, C);        // Y3 < 18p
    x1 = D;
    y1 = C;
}

// Co-Z addition with update (Meloni; Goundar-Joye-Miyaji "ZADDU"): for p1 = (x1, y1) and
// p2 = (x2, y2) sharing Z, (x3, y3) = p1 + p2 and (x1, y1) <- p1, both with Z' = Z mu,
// mu = x1 - x2 (returned).  p1 != +-p2.  2S + 4M.
// In: coordinates < 18 p.  Out: x1, y1 < 2 p; x3 < 6 p, y3 < 4 p; mu < 50 p.
FK20_HD void g1j_zaddu(fe &x3, fe &y3, fe &mu, fe &x1, fe &y1, const fe &x2, const fe &y2) {
    fe C, W1, W2, D, dy, t;
    fe_sub<32>(mu, x1, x2);         // < 50p
    fe_sqr(C, mu);                  // 2500p^2 -> < 2p
    fe_mul(W1, x1, C);              // < 2p
    fe_mul(W2, x2, C);              // < 2p
    fe_sub<32>(dy, y1, y2);         // < 50p
    fe_sqr(D, dy);                  // < 2p
    fe_sub<2>(t, W1, W2);           // < 4p
    fe_mul(y1, y1, t);              // A1 = Y1 (W1 - W2) < 2p
    fe_sub2<4>(x3, D, W1, W2);      // X3 = D - W1 - W2 < 6p
    fe_sub<8>(t, W1, x3);           // < 10p
    fe_mul(t, dy, t);               // 500p^2 -> < 2p
    fe_sub<2>(y3, t, y1);           // Y3 = dy (W1 - X3) - A1 < 4p
    x1 = W1;
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
