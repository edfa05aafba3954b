/*
 * fk20_oracle.c -- CPU restatement of the FK20 multi-proof path.
 *
 * THIS IS TEST INFRASTRUCTURE (the oracle), NOT PRODUCT CODE.  It is the parity reference
 * for the HIP library and the timed CPU baseline; the product never links it.
 *
 * Algorithm sources restated here (reference = lCardosoSantos/FK20-CUDA):
 *   - FK20 multi-proof, data-availability form: FK20Py/fk20_multi.py:64-104
 *     (fk20_multi_data_availability_optimized), FK20Py/fk20_single.py:45-84
 *     (toeplitz_part1/2/3), FK20Py/fft.py:17-55 (radix-2 FFT, inverse = reversed
 *     roots times n^-1), FK20Py/kzg_proofs.py:22-44 (generate_setup, root of unity
 *     from PRIMITIVE_ROOT = 5), kzg_proofs.py:198-205 (compute_proof_multi).
 *   - The CUDA stage split and layouts: fk20_poly2toeplitz_coefficients.cu:40-56,
 *     fk20_msm.cu:23-66, fk20_poly2h_fft.cu:27-80, fr_fft.cu:28-168 (natural-order
 *     DFT, bit-reversed load), g1p_fft.cu:25-170 and 180-330 (forward and inverse
 *     order of operations, which fixes the result on points outside G1), g1p_mul.cu:18-66 (MSB-first
 *     double-and-add over the raw 256-bit scalar), g1p.cu:95-115 (infinity and
 *     generator), g1a.cu:48-56 (affine conversion, affine infinity = (0,0)).
 *   - G1 formulas: complete projective formulas of Renes, Costello and Batina,
 *     eprint 2015/1060, algorithms 7 (add), 8 (mixed add) and 9 (double) for a = 0,
 *     b = 4, as the reference uses (g1p_add.cu:78, g1p_dbl.cu:28).
 *
 * Field elements are held internally in Montgomery form with 64-bit limbs; all
 * public functions take and return the reference's standard form.
 */
#define _GNU_SOURCE
#include "fk20_oracle.h"

#include <pthread.h>
#include <sched.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

typedef uint64_t u64;
typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------- */
/* Generic Montgomery arithmetic over n 64-bit limbs (n = 4 for Fr, 6 for Fp) */

typedef struct {
    int n;
    u64 m[6];     /* modulus */
    u64 minv;     /* -m^-1 mod 2^64 */
    u64 r1[6];    /* R mod m (Montgomery one) */
    u64 r2[6];    /* R^2 mod m */
} field_t;

static field_t FP, FR;

static const u64 P_LIMBS[6] = {
    0xb9feffffffffaaabULL, 0x1eabfffeb153ffffULL, 0x6730d2a0f6b0f624ULL,
    0x64774b84f38512bfULL, 0x4b1ba7b6434bacd7ULL, 0x1a0111ea397fe69aULL};
static const u64 R_LIMBS[4] = {
    0xffffffff00000001ULL, 0x53bda402fffe5bfeULL, 0x3339d80809a1d805ULL,
    0x73eda753299d7d48ULL};

static inline int mp_geq(const u64 *a, const u64 *b, int n) {
    for (int i = n - 1; i >= 0; i--) {
        if (a[i] != b[i]) return a[i] > b[i];
    }
    return 1;
}

static inline u64 mp_sub(u64 *z, const u64 *a, const u64 *b, int n) {
    u64 borrow = 0;
    for (int i = 0; i < n; i++) {
        u128 d = (u128)a[i] - b[i] - borrow;
        z[i] = (u64)d;
        borrow = (u64)(d >> 64) & 1;
    }
    return borrow;
}

static inline u64 mp_add(u64 *z, const u64 *a, const u64 *b, int n) {
    u64 carry = 0;
    for (int i = 0; i < n; i++) {
        u128 s = (u128)a[i] + b[i] + carry;
        z[i] = (u64)s;
        carry = (u64)(s >> 64);
    }
    return carry;
}

/* CIOS Montgomery product; a, b < m gives z < m.  Also valid for a < R, b < m. */
static inline void mont_mul(u64 *z, const u64 *a, const u64 *b, const field_t *F) {
    const int n = F->n;
    u64 t[8] = {0};
    for (int i = 0; i < n; i++) {
        u64 c = 0;
        for (int j = 0; j < n; j++) {
            u128 s = (u128)a[j] * b[i] + t[j] + c;
            t[j] = (u64)s;
            c = (u64)(s >> 64);
        }
        u128 s = (u128)t[n] + c;
        t[n] = (u64)s;
        t[n + 1] = (u64)(s >> 64);
        u64 mm = t[0] * F->minv;
        s = (u128)mm * F->m[0] + t[0];
        c = (u64)(s >> 64);
        for (int j = 1; j < n; j++) {
            s = (u128)mm * F->m[j] + t[j] + c;
            t[j - 1] = (u64)s;
            c = (u64)(s >> 64);
        }
        s = (u128)t[n] + c;
        t[n - 1] = (u64)s;
        t[n] = t[n + 1] + (u64)(s >> 64);
    }
    if (t[n] || mp_geq(t, F->m, n))
        mp_sub(t, t, F->m, n);
    memcpy(z, t, sizeof(u64) * n);
}

static inline void mod_add(u64 *z, const u64 *a, const u64 *b, const field_t *F) {
    u64 t[6];
    u64 c = mp_add(t, a, b, F->n);
    if (c || mp_geq(t, F->m, F->n))
        mp_sub(t, t, F->m, F->n);
    memcpy(z, t, sizeof(u64) * F->n);
}

static inline void mod_sub(u64 *z, const u64 *a, const u64 *b, const field_t *F) {
    u64 t[6];
    u64 br = mp_sub(t, a, b, F->n);
    if (br)
        mp_add(t, t, F->m, F->n);
    memcpy(z, t, sizeof(u64) * F->n);
}

static inline int mp_is_zero(const u64 *a, int n) {
    u64 acc = 0;
    for (int i = 0; i < n; i++) acc |= a[i];
    return acc == 0;
}

static void field_init(field_t *F, const u64 *m, int n) {
    memset(F, 0, sizeof(*F));
    F->n = n;
    memcpy(F->m, m, sizeof(u64) * n);
    u64 inv = 1;
    for (int i = 0; i < 7; i++) inv *= 2 - m[0] * inv;
    F->minv = (u64)0 - inv;
    /* R mod m and R^2 mod m by repeated doubling of 1 */
    u64 x[6] = {1, 0, 0, 0, 0, 0};
    for (int i = 0; i < 128 * n; i++) {
        u64 c = mp_add(x, x, x, n);
        if (c || mp_geq(x, m, n)) mp_sub(x, x, m, n);
        if (i == 64 * n - 1) memcpy(F->r1, x, sizeof(u64) * n);
    }
    memcpy(F->r2, x, sizeof(u64) * n);
}

static pthread_once_t init_once = PTHREAD_ONCE_INIT;
static void do_init(void) {
    field_init(&FP, P_LIMBS, 6);
    field_init(&FR, R_LIMBS, 4);
}
static inline void ensure_init(void) { pthread_once(&init_once, do_init); }

/* Any n-limb value (< 2^(64n)) to canonical Montgomery form. */
static inline void to_mont(u64 *z, const u64 *x, const field_t *F) { mont_mul(z, x, F->r2, F); }
static inline void from_mont(u64 *z, const u64 *x, const field_t *F) {
    u64 one[6] = {1, 0, 0, 0, 0, 0};
    mont_mul(z, x, one, F);
}

/* z = x^e (Montgomery in/out), e given as n-limb standard integer */
static void mont_pow(u64 *z, const u64 *x, const u64 *e, int elimbs, const field_t *F) {
    u64 acc[6], base[6];
    memcpy(acc, F->r1, sizeof(acc));
    memcpy(base, x, sizeof(u64) * F->n);
    for (int i = elimbs * 64 - 1; i >= 0; i--) {
        mont_mul(acc, acc, acc, F);
        if ((e[i / 64] >> (i % 64)) & 1) mont_mul(acc, acc, base, F);
    }
    memcpy(z, acc, sizeof(u64) * F->n);
}

static void mont_inv(u64 *z, const u64 *x, const field_t *F) {
    /* Fermat: x^(m-2); inverse of 0 is 0 (as the reference's fp_inv/fr_inv) */
    u64 e[6];
    u64 two[6] = {2, 0, 0, 0, 0, 0};
    mp_sub(e, F->m, two, F->n);
    mont_pow(z, x, e, F->n, F);
}

/* ------------------------------------------------------------------------- */
/* Fr / Fp public API (standard form) */

#define FIELD_API(PFX, F, N)                                                              \
    void or_##PFX##_reduce(u64 z[N], const u64 x[N]) {                                    \
        ensure_init();                                                                    \
        u64 t[6];                                                                         \
        to_mont(t, x, &F);                                                                \
        from_mont(z, t, &F);                                                              \
    }                                                                                     \
    void or_##PFX##_add(u64 z[N], const u64 x[N], const u64 y[N]) {                       \
        u64 a[6], b[6];                                                                   \
        or_##PFX##_reduce(a, x);                                                          \
        or_##PFX##_reduce(b, y);                                                          \
        mod_add(z, a, b, &F);                                                             \
    }                                                                                     \
    void or_##PFX##_sub(u64 z[N], const u64 x[N], const u64 y[N]) {                       \
        u64 a[6], b[6];                                                                   \
        or_##PFX##_reduce(a, x);                                                          \
        or_##PFX##_reduce(b, y);                                                          \
        mod_sub(z, a, b, &F);                                                             \
    }                                                                                     \
    void or_##PFX##_neg(u64 z[N], const u64 x[N]) {                                       \
        u64 zero[6] = {0};                                                                \
        or_##PFX##_sub(z, zero, x);                                                       \
    }                                                                                     \
    void or_##PFX##_mul(u64 z[N], const u64 x[N], const u64 y[N]) {                       \
        ensure_init();                                                                    \
        u64 a[6], b[6];                                                                   \
        to_mont(a, x, &F);                                                                \
        to_mont(b, y, &F);                                                                \
        mont_mul(a, a, b, &F);                                                            \
        from_mont(z, a, &F);                                                              \
    }                                                                                     \
    void or_##PFX##_sqr(u64 z[N], const u64 x[N]) { or_##PFX##_mul(z, x, x); }            \
    void or_##PFX##_inv(u64 z[N], const u64 x[N]) {                                       \
        ensure_init();                                                                    \
        u64 a[6];                                                                         \
        to_mont(a, x, &F);                                                                \
        mont_inv(a, a, &F);                                                               \
        from_mont(z, a, &F);                                                              \
    }                                                                                     \
    int or_##PFX##_eq(const u64 x[N], const u64 y[N]) {                                   \
        u64 a[6], b[6];                                                                   \
        or_##PFX##_reduce(a, x);                                                          \
        or_##PFX##_reduce(b, y);                                                          \
        return memcmp(a, b, sizeof(u64) * N) == 0;                                        \
    }

FIELD_API(fr, FR, 4)
FIELD_API(fp, FP, 6)

void or_fr_pow(u64 z[4], const u64 x[4], const u64 e[4]) {
    ensure_init();
    u64 a[4];
    to_mont(a, x, &FR);
    mont_pow(a, a, e, 4, &FR);
    from_mont(z, a, &FR);
}

/* kzg_proofs.py:39-44: root of unity of order n is 5^((r-1)/n) */
static void fr_root_mont(u64 z[4], u64 n) {
    ensure_init();
    /* (r-1)/n for n a power of two dividing 2^32: shift right */
    u64 e[4];
    u64 one[4] = {1, 0, 0, 0};
    mp_sub(e, R_LIMBS, one, 4);
    int sh = 0;
    while ((1ULL << sh) < n) sh++;
    for (int s = 0; s < sh; s++) {
        for (int i = 0; i < 4; i++) e[i] = (e[i] >> 1) | (i < 3 ? (e[i + 1] << 63) : 0);
    }
    u64 five[4] = {5, 0, 0, 0}, fm[4];
    to_mont(fm, five, &FR);
    mont_pow(z, fm, e, 4, &FR);
}

void or_fr_root_of_unity(u64 z[4], u64 n) {
    u64 t[4];
    fr_root_mont(t, n);
    from_mont(z, t, &FR);
}

/* fr_roots.cu:17-533: w^i for i = 0..512, then 2^-9 and 2^-9 * w^-128 */
void or_fr_roots515(u64 *out) {
    u64 w[4], acc[4];
    fr_root_mont(w, 512);
    memcpy(acc, FR.r1, sizeof(acc));
    for (int i = 0; i <= 512; i++) {
        from_mont(out + 4 * i, acc, &FR);
        mont_mul(acc, acc, w, &FR);
    }
    u64 v512[4] = {512, 0, 0, 0}, inv512[4], t[4], winv128[4];
    to_mont(t, v512, &FR);
    mont_inv(inv512, t, &FR);
    from_mont(out + 4 * 513, inv512, &FR);
    /* w^-128 = w^384 */
    u64 e384[4] = {384, 0, 0, 0};
    mont_pow(winv128, w, e384, 4, &FR);
    mont_mul(t, inv512, winv128, &FR);
    from_mont(out + 4 * 514, t, &FR);
}

/* ------------------------------------------------------------------------- */
/* G1 in Montgomery form */

typedef struct { u64 x[6], y[6], z[6]; } gp;

static inline void fp_m(u64 *z, const u64 *a, const u64 *b) { mont_mul(z, a, b, &FP); }
static inline void fp_a(u64 *z, const u64 *a, const u64 *b) { mod_add(z, a, b, &FP); }
static inline void fp_s(u64 *z, const u64 *a, const u64 *b) { mod_sub(z, a, b, &FP); }
/* z = 12 a (b3 = 3 * b with b = 4) */
static inline void fp_x12(u64 *z, const u64 *a) {
    u64 t[6];
    fp_a(t, a, a);  /* 2 */
    fp_a(t, t, a);  /* 3 */
    fp_a(t, t, t);  /* 6 */
    fp_a(z, t, t);  /* 12 */
}

static void gp_inf(gp *p) {
    memset(p, 0, sizeof(*p));
    memcpy(p->y, FP.r1, sizeof(p->y));
}

/* RCB 2015/1060 algorithm 7 (a = 0) */
static void gp_add(gp *r, const gp *p, const gp *q) {
    u64 t0[6], t1[6], t2[6], t3[6], t4[6], X3[6], Y3[6], Z3[6];
    const u64 *X1 = p->x, *Y1 = p->y, *Z1 = p->z, *X2 = q->x, *Y2 = q->y, *Z2 = q->z;
    fp_m(t0, X1, X2); fp_m(t1, Y1, Y2); fp_m(t2, Z1, Z2);
    fp_a(t3, X1, Y1); fp_a(t4, X2, Y2); fp_m(t3, t3, t4);
    fp_a(t4, t0, t1); fp_s(t3, t3, t4); fp_a(t4, Y1, Z1);
    fp_a(X3, Y2, Z2); fp_m(t4, t4, X3); fp_a(X3, t1, t2);
    fp_s(t4, t4, X3); fp_a(X3, X1, Z1); fp_a(Y3, X2, Z2);
    fp_m(X3, X3, Y3); fp_a(Y3, t0, t2); fp_s(Y3, X3, Y3);
    fp_a(X3, t0, t0); fp_a(t0, X3, t0); fp_x12(t2, t2);
    fp_a(Z3, t1, t2); fp_s(t1, t1, t2); fp_x12(Y3, Y3);
    fp_m(X3, t4, Y3); fp_m(t2, t3, t1); fp_s(X3, t2, X3);
    fp_m(Y3, Y3, t0); fp_m(t1, t1, Z3); fp_a(Y3, t1, Y3);
    fp_m(t0, t0, t3); fp_m(Z3, Z3, t4); fp_a(Z3, Z3, t0);
    memcpy(r->x, X3, 48); memcpy(r->y, Y3, 48); memcpy(r->z, Z3, 48);
}

/* RCB 2015/1060 algorithm 8 (mixed, a = 0); (X2, Y2) affine and not infinity */
static void gp_add_aff(gp *r, const gp *p, const u64 *X2, const u64 *Y2) {
    u64 t0[6], t1[6], t2[6], t3[6], t4[6], X3[6], Y3[6], Z3[6];
    const u64 *X1 = p->x, *Y1 = p->y, *Z1 = p->z;
    fp_m(t0, X1, X2); fp_m(t1, Y1, Y2); fp_a(t3, X2, Y2);
    fp_a(t4, X1, Y1); fp_m(t3, t3, t4); fp_a(t4, t0, t1);
    fp_s(t3, t3, t4); fp_m(t4, Y2, Z1); fp_a(t4, t4, Y1);
    fp_m(Y3, X2, Z1); fp_a(Y3, Y3, X1); fp_a(X3, t0, t0);
    fp_a(t0, X3, t0); fp_x12(t2, Z1); fp_a(Z3, t1, t2);
    fp_s(t1, t1, t2); fp_x12(Y3, Y3); fp_m(X3, t4, Y3);
    fp_m(t2, t3, t1); fp_s(X3, t2, X3); fp_m(Y3, Y3, t0);
    fp_m(t1, t1, Z3); fp_a(Y3, t1, Y3); fp_m(t0, t0, t3);
    fp_m(Z3, Z3, t4); fp_a(Z3, Z3, t0);
    memcpy(r->x, X3, 48); memcpy(r->y, Y3, 48); memcpy(r->z, Z3, 48);
}

/* RCB 2015/1060 algorithm 9 (a = 0) */
static void gp_dbl(gp *r, const gp *p) {
    u64 t0[6], t1[6], t2[6], X3[6], Y3[6], Z3[6];
    const u64 *X = p->x, *Y = p->y, *Z = p->z;
    fp_m(t0, Y, Y); fp_a(Z3, t0, t0); fp_a(Z3, Z3, Z3);
    fp_a(Z3, Z3, Z3); fp_m(t1, Y, Z); fp_m(t2, Z, Z);
    fp_x12(t2, t2); fp_m(X3, t2, Z3); fp_a(Y3, t0, t2);
    fp_m(Z3, t1, Z3); fp_a(t1, t2, t2); fp_a(t2, t1, t2);
    fp_s(t0, t0, t2); fp_m(Y3, t0, Y3); fp_a(Y3, X3, Y3);
    fp_m(t1, X, Y); fp_m(X3, t0, t1); fp_a(X3, X3, X3);
    memcpy(r->x, X3, 48); memcpy(r->y, Y3, 48); memcpy(r->z, Z3, 48);
}

static void gp_neg(gp *r, const gp *p) {
    u64 zero[6] = {0};
    *r = *p;
    fp_s(r->y, zero, p->y);
}

/* g1p_mul.cu:18-66: MSB-first double-and-add over the raw 256-bit scalar */
static void gp_mul(gp *r, const gp *p, const u64 k[4]) {
    gp q, base = *p;
    gp_inf(&q);
    int started = 0;
    for (int i = 255; i >= 0; i--) {
        int bit = (int)((k[i / 64] >> (i % 64)) & 1);
        if (started) gp_dbl(&q, &q);
        if (bit) {
            gp_add(&q, &q, &base);
            started = 1;
        }
    }
    *r = q;
}

static void gp_from_std(gp *r, const u64 *p) {
    to_mont(r->x, p, &FP);
    to_mont(r->y, p + 6, &FP);
    to_mont(r->z, p + 12, &FP);
}

static void gp_to_std(u64 *p, const gp *r) {
    from_mont(p, r->x, &FP);
    from_mont(p + 6, r->y, &FP);
    from_mont(p + 12, r->z, &FP);
}

static void gp_to_affine_m(u64 *ax, u64 *ay, const gp *p) {
    u64 zi[6];
    mont_inv(zi, p->z, &FP);
    fp_m(ax, p->x, zi);
    fp_m(ay, p->y, zi);
}

/* ------------------------------------------------------------------------- */
/* G1 public API */

void or_g1p_inf(u64 p[18]) {
    memset(p, 0, 18 * sizeof(u64));
    p[6] = 1;
}

void or_g1p_gen(u64 p[18]) {
    static const u64 gx[6] = {0xFB3AF00ADB22C6BBULL, 0x6C55E83FF97A1AEFULL, 0xA14E3A3F171BAC58ULL,
                              0xC3688C4F9774B905ULL, 0x2695638C4FA9AC0FULL, 0x17F1D3A73197D794ULL};
    static const u64 gy[6] = {0x0CAA232946C5E7E1ULL, 0xD03CC744A2888AE4ULL, 0x00DB18CB2C04B3EDULL,
                              0xFCF5E095D5D00AF6ULL, 0xA09E30ED741D8AE4ULL, 0x08B3F481E3AAA0F1ULL};
    memcpy(p, gx, 48);
    memcpy(p + 6, gy, 48);
    memset(p + 12, 0, 48);
    p[12] = 1;
}

int or_g1p_is_inf(const u64 p[18]) {
    ensure_init();
    gp a;
    gp_from_std(&a, p);
    return mp_is_zero(a.x, 6) && mp_is_zero(a.z, 6) && !mp_is_zero(a.y, 6);
}

/* Y^2 Z = X^3 + 4 Z^3, and infinity (0, y != 0, 0) */
int or_g1p_is_point(const u64 p[18]) {
    ensure_init();
    gp a;
    gp_from_std(&a, p);
    if (mp_is_zero(a.z, 6)) return mp_is_zero(a.x, 6) && !mp_is_zero(a.y, 6);
    u64 l[6], r[6], t[6], four[6] = {4, 0, 0, 0, 0, 0}, f[6];
    to_mont(f, four, &FP);
    fp_m(l, a.y, a.y); fp_m(l, l, a.z);
    fp_m(r, a.x, a.x); fp_m(r, r, a.x);
    fp_m(t, a.z, a.z); fp_m(t, t, a.z); fp_m(t, t, f);
    fp_a(r, r, t);
    return memcmp(l, r, 48) == 0;
}

/* Projective equality by cross-multiplication (as g1p_compare.cu:23-54) */
int or_g1p_eq(const u64 p[18], const u64 q[18]) {
    ensure_init();
    gp a, b;
    gp_from_std(&a, p);
    gp_from_std(&b, q);
    int ia = mp_is_zero(a.z, 6), ib = mp_is_zero(b.z, 6);
    if (ia || ib) return ia && ib;
    u64 l[6], r[6];
    fp_m(l, a.x, b.z); fp_m(r, b.x, a.z);
    if (memcmp(l, r, 48)) return 0;
    fp_m(l, a.y, b.z); fp_m(r, b.y, a.z);
    return memcmp(l, r, 48) == 0;
}

void or_g1p_add(u64 r[18], const u64 p[18], const u64 q[18]) {
    ensure_init();
    gp a, b;
    gp_from_std(&a, p);
    gp_from_std(&b, q);
    gp_add(&a, &a, &b);
    gp_to_std(r, &a);
}

void or_g1p_sub(u64 r[18], const u64 p[18], const u64 q[18]) {
    ensure_init();
    gp a, b;
    gp_from_std(&a, p);
    gp_from_std(&b, q);
    gp_neg(&b, &b);
    gp_add(&a, &a, &b);
    gp_to_std(r, &a);
}

void or_g1p_dbl(u64 r[18], const u64 p[18]) {
    ensure_init();
    gp a;
    gp_from_std(&a, p);
    gp_dbl(&a, &a);
    gp_to_std(r, &a);
}

void or_g1p_neg(u64 r[18], const u64 p[18]) {
    ensure_init();
    gp a;
    gp_from_std(&a, p);
    gp_neg(&a, &a);
    gp_to_std(r, &a);
}

void or_g1p_add_affine(u64 r[18], const u64 p[18], const u64 q[12]) {
    ensure_init();
    gp a;
    gp_from_std(&a, p);
    u64 x[6], y[6];
    to_mont(x, q, &FP);
    to_mont(y, q + 6, &FP);
    if (!(mp_is_zero(x, 6) && mp_is_zero(y, 6)))  /* affine infinity is (0,0) */
        gp_add_aff(&a, &a, x, y);
    gp_to_std(r, &a);
}

void or_g1p_mul(u64 r[18], const u64 p[18], const u64 k[4]) {
    ensure_init();
    gp a;
    gp_from_std(&a, p);
    gp_mul(&a, &a, k);
    gp_to_std(r, &a);
}

void or_g1p_to_affine(u64 a[12], const u64 p[18]) {
    ensure_init();
    gp m;
    gp_from_std(&m, p);
    if (mp_is_zero(m.z, 6)) {
        memset(a, 0, 12 * sizeof(u64));
        return;
    }
    u64 x[6], y[6];
    gp_to_affine_m(x, y, &m);
    from_mont(a, x, &FP);
    from_mont(a + 6, y, &FP);
}

void or_g1p_to_affine_n(u64 *a, const u64 *p, int64_t count) {
    for (int64_t i = 0; i < count; i++) or_g1p_to_affine(a + 12 * i, p + 18 * i);
}

void or_g1p_normalize(u64 r[18], const u64 p[18]) {
    u64 a[12];
    or_g1p_to_affine(a, p);
    if (mp_is_zero(a, 12)) {
        or_g1p_inf(r);
        return;
    }
    memcpy(r, a, 96);
    memset(r + 12, 0, 48);
    r[12] = 1;
}

/* ------------------------------------------------------------------------- */
/* Transforms */

static unsigned bitrev(unsigned x, int bits) {
    unsigned r = 0;
    for (int i = 0; i < bits; i++) r |= ((x >> i) & 1) << (bits - 1 - i);
    return r;
}

static int log2i(int n) {
    int b = 0;
    while ((1 << b) < n) b++;
    return b;
}

/* Montgomery powers w^0..w^(n-1) of the order-n root (or its inverse) */
static u64 *fr_powers_mont(int n, int inverse) {
    u64 *tw = (u64 *)malloc(sizeof(u64) * 4 * (size_t)n);
    u64 w[4];
    fr_root_mont(w, (u64)n);
    if (inverse) mont_inv(w, w, &FR);
    memcpy(tw, FR.r1, 32);
    for (int i = 1; i < n; i++) mont_mul(tw + 4 * i, tw + 4 * (i - 1), w, &FR);
    return tw;
}

/* In-place radix-2 DIT on Montgomery Fr values (fr_fft.cu:28-168 structure) */
static void fr_fft_mont(u64 *a, int n, int inverse) {
    int bits = log2i(n);
    u64 *tw = fr_powers_mont(n, inverse);
    for (int i = 0; i < n; i++) {
        int j = (int)bitrev((unsigned)i, bits);
        if (j > i) {
            u64 t[4];
            memcpy(t, a + 4 * i, 32);
            memcpy(a + 4 * i, a + 4 * j, 32);
            memcpy(a + 4 * j, t, 32);
        }
    }
    for (int half = 1; half < n; half *= 2) {
        int step = n / (2 * half);
        for (int blk = 0; blk < n; blk += 2 * half) {
            for (int t = 0; t < half; t++) {
                u64 *l = a + 4 * (blk + t), *r = a + 4 * (blk + t + half), x[4];
                mont_mul(x, r, tw + 4 * (t * step), &FR);
                mod_sub(r, l, x, &FR);
                mod_add(l, l, x, &FR);
            }
        }
    }
    if (inverse) {
        u64 nn[4] = {(u64)n, 0, 0, 0}, ninv[4];
        to_mont(ninv, nn, &FR);
        mont_inv(ninv, ninv, &FR);
        for (int i = 0; i < n; i++) mont_mul(a + 4 * i, a + 4 * i, ninv, &FR);
    }
    free(tw);
}

void or_fr_fft(u64 *out, const u64 *in, int n, int inverse) {
    ensure_init();
    u64 *a = (u64 *)malloc(sizeof(u64) * 4 * (size_t)n);
    for (int i = 0; i < n; i++) to_mont(a + 4 * i, in + 4 * i, &FR);
    fr_fft_mont(a, n, inverse);
    for (int i = 0; i < n; i++) from_mont(out + 4 * i, a + 4 * i, &FR);
    free(a);
}

/* Standard-form twiddles w^i for the G1 butterflies (scalar multiplier input) */
static u64 *fr_powers_std(int n, int inverse) {
    u64 *tw = fr_powers_mont(n, inverse);
    for (int i = 0; i < n; i++) from_mont(tw + 4 * i, tw + 4 * i, &FR);
    return tw;
}

/* In-place G1 inverse DFT in the reference's order of operations (g1p_fft.cu:180-330):
 * natural-order input; stages s = log2(n)-1 .. 0, each butterfly (l, r) <- (l + r, l - r)
 * followed, for s >= 2, by r <- [w^-(t * step)] r; stage 1 instead multiplies l by 2^-log2(n)
 * and r by 2^-log2(n) (t = 0) or 2^-log2(n) w^-(n/4) (t = 1) (fr_roots[513], fr_roots[514]);
 * output in natural order from the bit-reversed positions.  On G1 this is the DFT with the
 * inverse roots times n^-1 (FK20Py/fft.py:39-55); on points outside G1, where [a]P depends on
 * the integer a and not only on a mod r, only this exact sequence of canonical multipliers
 * gives the reference's result. */
static void g1_ift_m(gp *a, int n) {
    int bits = log2i(n);
    u64 *tw = fr_powers_std(n, 1);
    u64 nn[4] = {(u64)n, 0, 0, 0}, ninv[4], ninv_std[4], ninv_w_std[4], t4[4];
    to_mont(ninv, nn, &FR);
    mont_inv(ninv, ninv, &FR);
    from_mont(ninv_std, ninv, &FR);
    to_mont(t4, tw + 4 * (n / 4), &FR);
    mont_mul(t4, t4, ninv, &FR);
    from_mont(ninv_w_std, t4, &FR);
    for (int half = n / 2; half >= 1; half /= 2) {
        int step = n / (2 * half);
        for (int blk = 0; blk < n; blk += 2 * half) {
            for (int t = 0; t < half; t++) {
                gp *l = a + blk + t, *r = a + blk + t + half, nr, sum, dif;
                gp_neg(&nr, r);
                gp_add(&sum, l, r);
                gp_add(&dif, l, &nr);
                if (half == 2) {
                    gp_mul(&sum, &sum, ninv_std);
                    gp_mul(&dif, &dif, t ? ninv_w_std : ninv_std);
                } else if (half > 2 && t) {
                    gp_mul(&dif, &dif, tw + 4 * (t * step));
                }
                *l = sum;
                *r = dif;
            }
        }
    }
    for (int i = 0; i < n; i++) {
        int j = (int)bitrev((unsigned)i, bits);
        if (j > i) {
            gp t = a[i];
            a[i] = a[j];
            a[j] = t;
        }
    }
    free(tw);
}

/* In-place G1 DFT (g1p_fft.cu:25-170 structure; FK20Py/fft.py:39-55 semantics); the inverse
 * is g1_ift_m */
static void g1_fft_m(gp *a, int n, int inverse) {
    if (inverse) {
        g1_ift_m(a, n);
        return;
    }
    int bits = log2i(n);
    u64 *tw = fr_powers_std(n, 0);
    for (int i = 0; i < n; i++) {
        int j = (int)bitrev((unsigned)i, bits);
        if (j > i) {
            gp t = a[i];
            a[i] = a[j];
            a[j] = t;
        }
    }
    for (int half = 1; half < n; half *= 2) {
        int step = n / (2 * half);
        for (int blk = 0; blk < n; blk += 2 * half) {
            for (int t = 0; t < half; t++) {
                gp *l = a + blk + t, *r = a + blk + t + half, x, nx;
                if (t == 0) x = *r;
                else gp_mul(&x, r, tw + 4 * (t * step));
                gp_neg(&nx, &x);
                gp_add(r, l, &nx);
                gp_add(l, l, &x);
            }
        }
    }
    free(tw);
}

void or_g1p_fft(u64 *out, const u64 *in, int n, int inverse) {
    ensure_init();
    gp *a = (gp *)malloc(sizeof(gp) * (size_t)n);
    for (int i = 0; i < n; i++) gp_from_std(a + i, in + 18 * i);
    g1_fft_m(a, n, inverse);
    for (int i = 0; i < n; i++) gp_to_std(out + 18 * i, a + i);
    free(a);
}

/* ------------------------------------------------------------------------- */
/* FK20 stages */

/* fk20_multi.py:92-93 (reduced_polynomial[-i-1::l] + [0]*(k+1) +
 * reduced_polynomial[2l-i-1:-l-i:l]); same layout as fk20_poly2toeplitz_coefficients.cu:40-56 */
static void toeplitz_row_std(u64 *tc, const u64 *poly, int N, int i) {
    const int l = 16, k = N / l, n2 = 2 * k;
    memset(tc, 0, sizeof(u64) * 4 * (size_t)n2);
    memcpy(tc, poly + 4 * (N - 1 - i), 32);
    int pos = k + 2;
    for (int src = 2 * l - i - 1; src < N - l - i; src += l) memcpy(tc + 4 * (pos++), poly + 4 * src, 32);
}

void or_fk20_poly2toeplitz_coefficients(u64 *tc, const u64 *poly, int N) {
    const int n2 = N / 8;
    for (int i = 0; i < 16; i++) toeplitz_row_std(tc + 4 * (size_t)n2 * i, poly, N, i);
}

/* fk20_msm.cu:23-66, FK20Py toeplitz_part2 summed over i (fk20_multi.py:95-96) */
static void msm_m(gp *hext, const u64 *tc_fft_std, const gp *xext, int n2) {
    for (int j = 0; j < n2; j++) {
        gp acc, t;
        gp_inf(&acc);
        for (int i = 0; i < 16; i++) {
            gp_mul(&t, xext + (size_t)n2 * i + j, tc_fft_std + 4 * ((size_t)n2 * i + j));
            gp_add(&acc, &acc, &t);
        }
        hext[j] = acc;
    }
}

void or_fk20_msm(u64 *hext_fft, const u64 *tc_fft, const u64 *xext_fft, int n2) {
    ensure_init();
    gp *x = (gp *)malloc(sizeof(gp) * 16 * (size_t)n2);
    gp *h = (gp *)malloc(sizeof(gp) * (size_t)n2);
    for (int i = 0; i < 16 * n2; i++) gp_from_std(x + i, xext_fft + 18 * (size_t)i);
    msm_m(h, tc_fft, x, n2);
    for (int j = 0; j < n2; j++) gp_to_std(hext_fft + 18 * (size_t)j, h + j);
    free(x);
    free(h);
}

/* fk20_single.py:87-93 (toeplitz_part3) + fk20_multi.py:98-104: ifft, keep first half,
 * pad with infinity, fft */
static void hext2hfft_m(gp *a, int n2) {
    g1_fft_m(a, n2, 1);
    for (int j = n2 / 2; j < n2; j++) gp_inf(a + j);
    g1_fft_m(a, n2, 0);
}

void or_fk20_hext_fft2h_fft(u64 *h_fft, const u64 *hext_fft, int n2) {
    ensure_init();
    gp *a = (gp *)malloc(sizeof(gp) * (size_t)n2);
    for (int j = 0; j < n2; j++) gp_from_std(a + j, hext_fft + 18 * (size_t)j);
    hext2hfft_m(a, n2);
    for (int j = 0; j < n2; j++) gp_to_std(h_fft + 18 * (size_t)j, a + j);
    free(a);
}

static void fk20_row_m(gp *hfft, const u64 *poly, const gp *xext, int N) {
    const int n2 = N / 8;
    u64 *tc = (u64 *)malloc(sizeof(u64) * 4 * 16 * (size_t)n2);
    or_fk20_poly2toeplitz_coefficients(tc, poly, N);
    for (int i = 0; i < 16; i++) or_fr_fft(tc + 4 * (size_t)n2 * i, tc + 4 * (size_t)n2 * i, n2, 0);
    msm_m(hfft, tc, xext, n2);
    hext2hfft_m(hfft, n2);
    free(tc);
}

typedef struct {
    u64 *h_fft;
    const u64 *poly;
    const gp *xext;
    int N, row0, row1;
} row_job;

static void *row_worker(void *arg) {
    row_job *jb = (row_job *)arg;
    const int n2 = jb->N / 8;
    gp *h = (gp *)malloc(sizeof(gp) * (size_t)n2);
    for (int r = jb->row0; r < jb->row1; r++) {
        fk20_row_m(h, jb->poly + 4 * (size_t)jb->N * r, jb->xext, jb->N);
        for (int j = 0; j < n2; j++) gp_to_std(jb->h_fft + 18 * ((size_t)n2 * r + j), h + j);
    }
    free(h);
    return NULL;
}

int or_default_threads(void) {
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (int)n : 1;
}

void or_fk20_poly2h_fft(u64 *h_fft, const u64 *poly, const u64 *xext_fft, int N, int rows,
                        int nthreads) {
    ensure_init();
    const int n2 = N / 8;
    gp *x = (gp *)malloc(sizeof(gp) * 16 * (size_t)n2);
    for (int i = 0; i < 16 * n2; i++) gp_from_std(x + i, xext_fft + 18 * (size_t)i);
    if (nthreads <= 0) nthreads = or_default_threads();
    if (nthreads > rows) nthreads = rows;
    if (nthreads < 1) nthreads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    row_job *jobs = (row_job *)malloc(sizeof(row_job) * (size_t)nthreads);
    for (int t = 0; t < nthreads; t++) {
        jobs[t].h_fft = h_fft;
        jobs[t].poly = poly;
        jobs[t].xext = x;
        jobs[t].N = N;
        jobs[t].row0 = (int)((long)rows * t / nthreads);
        jobs[t].row1 = (int)((long)rows * (t + 1) / nthreads);
        if (nthreads == 1) row_worker(&jobs[t]);
        else pthread_create(&th[t], NULL, row_worker, &jobs[t]);
    }
    if (nthreads > 1)
        for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(th);
    free(jobs);
    free(x);
}

/* ------------------------------------------------------------------------- */
/* Setup */

/* kzg_proofs.py:22-28 generate_setup: [tau^i] G1, normalised to z = 1 */
void or_setup_srs(u64 *srs, const u64 tau[4], int count) {
    ensure_init();
    u64 g[18], tm[4], acc[4], accs[4];
    or_g1p_gen(g);
    gp G;
    gp_from_std(&G, g);
    to_mont(tm, tau, &FR);
    memcpy(acc, FR.r1, 32);
    for (int i = 0; i < count; i++) {
        gp t;
        from_mont(accs, acc, &FR);
        gp_mul(&t, &G, accs);
        u64 s[18];
        gp_to_std(s, &t);
        or_g1p_normalize(srs + 18 * (size_t)i, s);
        mont_mul(acc, acc, tm, &FR);
    }
}

/* KZG commitment of each row: commit[r] = sum_i poly[r][i] * srs[i] (FK20Py kzg_proofs.py:159
 * commit_to_poly = lincomb(setup[0][:len(polynomial)], polynomial)), by double-and-add. */
void or_commit(u64 *commit, const u64 *poly, const u64 *srs, int N, int rows) {
    ensure_init();
    for (int r = 0; r < rows; r++) {
        gp acc;
        gp_inf(&acc);
        for (int i = 0; i < N; i++) {
            gp base, t;
            gp_from_std(&base, srs + 18 * (size_t)i);
            gp_mul(&t, &base, poly + 4 * ((size_t)r * N + i));
            gp_add(&acc, &acc, &t);
        }
        gp_to_std(commit + 18 * (size_t)r, &acc);
    }
}

/* fk20_multi.py:82-85 + fk20_single.py:45-58 (toeplitz_part1) */
void or_setup2xext_fft(u64 *xext_fft, const u64 *srs, int N) {
    ensure_init();
    const int l = 16, k = N / l, n2 = 2 * k;
    gp *a = (gp *)malloc(sizeof(gp) * (size_t)n2);
    for (int i = 0; i < l; i++) {
        int m = 0;
        for (int src = N - l - 1 - i; src >= 0; src -= l) gp_from_std(a + (m++), srs + 18 * (size_t)src);
        for (; m < n2; m++) gp_inf(a + m);
        g1_fft_m(a, n2, 0);
        for (int j = 0; j < n2; j++) gp_to_std(xext_fft + 18 * ((size_t)n2 * i + j), a + j);
    }
    free(a);
}

void or_g1_gen_mul_affine_n(u64 *aff, const u64 *scalars, int64_t count, int nthreads);

/* Scalar shadow of xext_fft: xhat_i[m] = tau^(N-l-1-i-l*m) for m <= k-2, else 0; FFT in Fr */
static void xhat_fft_mont(u64 *xh, const u64 tau[4], int N) {
    const int l = 16, k = N / l, n2 = 2 * k;
    u64 tm[4];
    to_mont(tm, tau, &FR);
    u64 *pw = (u64 *)malloc(sizeof(u64) * 4 * (size_t)N);
    memcpy(pw, FR.r1, 32);
    for (int e = 1; e < N; e++) mont_mul(pw + 4 * e, pw + 4 * (e - 1), tm, &FR);
    for (int i = 0; i < l; i++) {
        u64 *row = xh + 4 * (size_t)n2 * i;
        memset(row, 0, sizeof(u64) * 4 * (size_t)n2);
        int m = 0;
        for (int src = N - l - 1 - i; src >= 0; src -= l) memcpy(row + 4 * (m++), pw + 4 * src, 32);
        fr_fft_mont(row, n2, 0);
    }
    free(pw);
}

void or_xext_fft_from_tau(u64 *xext_fft, const u64 tau[4], int N) {
    ensure_init();
    const int n2 = N / 8;
    u64 *xh = (u64 *)malloc(sizeof(u64) * 4 * 16 * (size_t)n2);
    u64 *aff = (u64 *)malloc(sizeof(u64) * 12 * 16 * (size_t)n2);
    xhat_fft_mont(xh, tau, N);
    for (int i = 0; i < 16 * n2; i++) from_mont(xh + 4 * (size_t)i, xh + 4 * (size_t)i, &FR);
    or_g1_gen_mul_affine_n(aff, xh, 16 * (int64_t)n2, 0);
    for (int i = 0; i < 16 * n2; i++) {
        u64 *p = xext_fft + 18 * (size_t)i;
        const u64 *a = aff + 12 * (size_t)i;
        if (mp_is_zero(a, 12)) {
            or_g1p_inf(p);
            continue;
        }
        memcpy(p, a, 96);
        memset(p + 12, 0, 48);
        p[12] = 1;
    }
    free(xh);
    free(aff);
}

/* ------------------------------------------------------------------------- */
/* Closed forms */

void or_fk20_shadow(u64 *s_fft, const u64 *poly, const u64 tau[4], int N) {
    ensure_init();
    const int k = N / 16, n2 = 2 * k;
    u64 *xh = (u64 *)malloc(sizeof(u64) * 4 * 16 * (size_t)n2);
    u64 *tc = (u64 *)malloc(sizeof(u64) * 4 * (size_t)n2);
    u64 *h = (u64 *)calloc((size_t)n2 * 4, sizeof(u64));
    xhat_fft_mont(xh, tau, N);
    for (int i = 0; i < 16; i++) {
        toeplitz_row_std(tc, poly, N, i);
        for (int j = 0; j < n2; j++) to_mont(tc + 4 * j, tc + 4 * j, &FR);
        fr_fft_mont(tc, n2, 0);
        for (int j = 0; j < n2; j++) {
            u64 t[4];
            mont_mul(t, tc + 4 * j, xh + 4 * ((size_t)n2 * i + j), &FR);
            mod_add(h + 4 * j, h + 4 * j, t, &FR);
        }
    }
    fr_fft_mont(h, n2, 1);
    memset(h + 4 * (size_t)k, 0, sizeof(u64) * 4 * (size_t)k);
    fr_fft_mont(h, n2, 0);
    for (int j = 0; j < n2; j++) from_mont(s_fft + 4 * j, h + 4 * j, &FR);
    free(xh);
    free(tc);
    free(h);
}

/* kzg_proofs.py:198-205: quotient of f by X^16 - c, evaluated at tau */
void or_kzg_quotient_at(u64 out[4], const u64 *poly, int N, const u64 c[4], const u64 tau[4]) {
    ensure_init();
    u64 *rem = (u64 *)malloc(sizeof(u64) * 4 * (size_t)N);
    u64 *q = (u64 *)calloc((size_t)N * 4, sizeof(u64));
    u64 cm[4], tm[4];
    to_mont(cm, c, &FR);
    to_mont(tm, tau, &FR);
    for (int i = 0; i < N; i++) to_mont(rem + 4 * i, poly + 4 * i, &FR);
    for (int d = N - 1; d >= 16; d--) {
        u64 t[4];
        memcpy(q + 4 * (d - 16), rem + 4 * d, 32);
        mont_mul(t, rem + 4 * d, cm, &FR);
        mod_add(rem + 4 * (d - 16), rem + 4 * (d - 16), t, &FR);
    }
    u64 acc[4] = {0, 0, 0, 0};
    for (int d = N - 17; d >= 0; d--) {
        mont_mul(acc, acc, tm, &FR);
        mod_add(acc, acc, q + 4 * d, &FR);
    }
    from_mont(out, acc, &FR);
    free(rem);
    free(q);
}

/* ------------------------------------------------------------------------- */
/* Bulk closed forms: [s]G for many scalars by fixed-base windows over G.
 *
 * The scalar-shadow closed form of the FK20 path (h_fft[k] = [s_k]G for a setup with known
 * secret) needs one multiplication of the generator per output point; these helpers do that
 * for whole batches.  The scalar is reduced mod r and split into 32 unsigned 8-bit windows;
 * gen_tab[w][d-1] = d 2^(8w) G (affine, Montgomery), so [s]G is at most 32 mixed additions
 * (RCB algorithm 8, complete for any accumulator).  The table is built once, by doublings
 * and additions over G, and converted to affine one point at a time. */

#define GEN_WIN 8
#define GEN_NWIN 32
#define GEN_H 255

static u64 (*gen_tab)[GEN_H][12];  /* [GEN_NWIN][GEN_H] x[6] y[6], Montgomery */
static pthread_once_t gen_once = PTHREAD_ONCE_INIT;

static void gen_tab_init(void) {
    ensure_init();
    gen_tab = malloc(sizeof(u64) * 12 * GEN_H * GEN_NWIN);
    u64 g[18];
    or_g1p_gen(g);
    gp base;
    gp_from_std(&base, g);
    for (int w = 0; w < GEN_NWIN; w++) {
        gp acc = base;
        for (int d = 1; d <= GEN_H; d++) {
            if (d > 1) gp_add(&acc, &acc, &base);
            gp_to_affine_m(gen_tab[w][d - 1], gen_tab[w][d - 1] + 6, &acc);
        }
        for (int b = 0; b < GEN_WIN; b++) gp_dbl(&base, &base);
    }
}

/* [s]G in Montgomery projective form, s any 256-bit value */
static void gen_mul_m(gp *r, const u64 s[4]) {
    pthread_once(&gen_once, gen_tab_init);
    u64 t[4], k[4];
    to_mont(t, s, &FR);
    from_mont(k, t, &FR);  /* k = s mod r */
    gp_inf(r);
    for (int w = 0; w < GEN_NWIN; w++) {
        const unsigned d = (unsigned)(k[w / 8] >> (8 * (w % 8))) & 0xffu;
        if (d) gp_add_aff(r, r, gen_tab[w][d - 1], gen_tab[w][d - 1] + 6);
    }
}

/* Canonical affine of n Montgomery points by Montgomery's batch inversion (one Fermat
 * inversion for the whole run); infinity -> (0, 0). */
static void gp_to_affine_batch(u64 *aff, const gp *p, int64_t n) {
    u64 *pre = malloc(sizeof(u64) * 6 * (size_t)(n + 1));
    memcpy(pre, FP.r1, 48);
    for (int64_t i = 0; i < n; i++) {
        const u64 *z = mp_is_zero(p[i].z, 6) ? FP.r1 : p[i].z;
        fp_m(pre + 6 * (i + 1), pre + 6 * i, z);
    }
    u64 inv[6];
    mont_inv(inv, pre + 6 * n, &FP);
    for (int64_t i = n - 1; i >= 0; i--) {
        u64 zi[6], x[6], y[6];
        const int inf = mp_is_zero(p[i].z, 6);
        fp_m(zi, inv, pre + 6 * i);
        if (!inf) fp_m(inv, inv, p[i].z);
        if (inf) {
            memset(aff + 12 * i, 0, 96);
            continue;
        }
        fp_m(x, p[i].x, zi);
        fp_m(y, p[i].y, zi);
        from_mont(aff + 12 * i, x, &FP);
        from_mont(aff + 12 * i + 6, y, &FP);
    }
    free(pre);
}

typedef struct {
    int kind;  /* 0: gen_mul affine, 1: gen_mul projective times z, 2: match, 3: shadow rows,
                  4: MSM scalar shadow */
    u64 *out;
    const u64 *a, *b;
    const u64 *tau;
    int N;
    int64_t i0, i1;
    int64_t bad, first;
} bulk_job;

static void *bulk_worker(void *arg) {
    bulk_job *jb = (bulk_job *)arg;
    const int64_t n = jb->i1 - jb->i0;
    if (jb->kind == 0) {
        gp *tmp = malloc(sizeof(gp) * (size_t)(n > 0 ? n : 1));
        for (int64_t i = 0; i < n; i++) gen_mul_m(tmp + i, jb->a + 4 * (jb->i0 + i));
        gp_to_affine_batch(jb->out + 12 * jb->i0, tmp, n);
        free(tmp);
    } else if (jb->kind == 1) {
        for (int64_t i = jb->i0; i < jb->i1; i++) {
            gp t;
            u64 zm[6];
            gen_mul_m(&t, jb->a + 4 * i);
            to_mont(zm, jb->b + 6 * i, &FP);
            if (mp_is_zero(t.z, 6)) {
                /* infinity as (0, z, 0): a non-normalised representative */
                memset(t.x, 0, 48);
                memcpy(t.y, zm, 48);
            } else {
                fp_m(t.x, t.x, zm);
                fp_m(t.y, t.y, zm);
                fp_m(t.z, t.z, zm);
            }
            gp_to_std(jb->out + 18 * i, &t);
        }
    } else if (jb->kind == 2) {
        jb->bad = 0;
        jb->first = -1;
        for (int64_t i = jb->i0; i < jb->i1; i++) {
            gp p;
            u64 ax[6], ay[6], l[6];
            gp_from_std(&p, jb->a + 18 * i);
            const u64 *e = jb->b + 12 * i;
            int ok;
            if (mp_is_zero(e, 12)) {
                ok = mp_is_zero(p.z, 6) && mp_is_zero(p.x, 6) && !mp_is_zero(p.y, 6);
            } else {
                to_mont(ax, e, &FP);
                to_mont(ay, e + 6, &FP);
                ok = !mp_is_zero(p.z, 6);
                fp_m(l, ax, p.z);
                ok = ok && memcmp(l, p.x, 48) == 0;
                fp_m(l, ay, p.z);
                ok = ok && memcmp(l, p.y, 48) == 0;
            }
            if (!ok) {
                if (jb->first < 0) jb->first = i;
                jb->bad++;
            }
        }
    } else if (jb->kind == 4) {
        /* out[r][j] = sum_i tc_fft[r][i][j] * xs[i][j] (fk20_msm.cu:23-66 on scalars) */
        const int n2 = jb->N;
        for (int64_t r = jb->i0; r < jb->i1; r++)
            for (int j = 0; j < n2; j++) {
                u64 acc[4] = {0, 0, 0, 0};
                for (int i = 0; i < 16; i++) {
                    u64 a[4], b[4];
                    to_mont(a, jb->a + 4 * (((size_t)r * 16 + i) * n2 + j), &FR);
                    to_mont(b, jb->b + 4 * ((size_t)i * n2 + j), &FR);
                    mont_mul(a, a, b, &FR);
                    mod_add(acc, acc, a, &FR);
                }
                from_mont(jb->out + 4 * ((size_t)r * n2 + j), acc, &FR);
            }
    } else {
        const int n2 = jb->N / 8;
        for (int64_t r = jb->i0; r < jb->i1; r++)
            or_fk20_shadow(jb->out + 4 * (size_t)n2 * r, jb->a + 4 * (size_t)jb->N * r, jb->tau, jb->N);
    }
    return NULL;
}

/* Default worker count of the bulk helpers: the CPUs this process may run on, at most 16 (the
 * per-GPU CPU share of a GPU box, whatever the machine's total). */
static int bulk_threads(void) {
    cpu_set_t set;
    int n = 0;
    if (sched_getaffinity(0, sizeof set, &set) == 0) n = CPU_COUNT(&set);
    if (n <= 0) n = or_default_threads();
    return n < 16 ? n : 16;
}

/* Splits [0, count) over nthreads workers of one kind; returns summed mismatches (kind 2). */
static int64_t bulk_run(bulk_job proto, int64_t count, int nthreads, int64_t *first) {
    ensure_init();
    pthread_once(&gen_once, gen_tab_init);
    if (nthreads <= 0) nthreads = bulk_threads();
    if (nthreads > count) nthreads = (int)(count > 0 ? count : 1);
    pthread_t *th = malloc(sizeof(pthread_t) * (size_t)nthreads);
    bulk_job *jobs = malloc(sizeof(bulk_job) * (size_t)nthreads);
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = proto;
        jobs[t].i0 = count * t / nthreads;
        jobs[t].i1 = count * (t + 1) / nthreads;
        if (nthreads == 1) bulk_worker(&jobs[t]);
        else pthread_create(&th[t], NULL, bulk_worker, &jobs[t]);
    }
    if (nthreads > 1)
        for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    int64_t bad = 0, f = -1;
    for (int t = 0; t < nthreads; t++) {
        bad += jobs[t].bad;
        if (f < 0 && jobs[t].first >= 0) f = jobs[t].first;
    }
    if (first) *first = f;
    free(th);
    free(jobs);
    return bad;
}

void or_g1_gen_mul_affine_n(u64 *aff, const u64 *scalars, int64_t count, int nthreads) {
    bulk_job j;
    memset(&j, 0, sizeof j);
    j.kind = 0;
    j.out = aff;
    j.a = scalars;
    bulk_run(j, count, nthreads, NULL);
}

void or_g1_gen_mul_proj_n(u64 *proj, const u64 *scalars, const u64 *zs, int64_t count, int nthreads) {
    bulk_job j;
    memset(&j, 0, sizeof j);
    j.kind = 1;
    j.out = proj;
    j.a = scalars;
    j.b = zs;
    bulk_run(j, count, nthreads, NULL);
}

int64_t or_g1p_match_affine_n(const u64 *p, const u64 *aff, int64_t count, int nthreads, int64_t *first) {
    bulk_job j;
    memset(&j, 0, sizeof j);
    j.kind = 2;
    j.a = p;
    j.b = aff;
    return bulk_run(j, count, nthreads, first);
}

void or_fk20_shadow_rows(u64 *s_fft, const u64 *poly, const u64 tau[4], int N, int rows, int nthreads) {
    bulk_job j;
    memset(&j, 0, sizeof j);
    j.kind = 3;
    j.out = s_fft;
    j.a = poly;
    j.tau = tau;
    j.N = N;
    bulk_run(j, rows, nthreads, NULL);
}

void or_fk20_msm_shadow(u64 *out, const u64 *tc_fft, const u64 *xs, int n2, int rows, int nthreads) {
    bulk_job j;
    memset(&j, 0, sizeof j);
    j.kind = 4;
    j.out = out;
    j.a = tc_fft;
    j.b = xs;
    j.N = n2;
    bulk_run(j, rows, nthreads, NULL);
}
