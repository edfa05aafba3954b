/*
 * fk20_oracle.h -- CPU restatement of the FK20 multi-proof path (test infrastructure).
 *
 * THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load it, and only as
 * the checker (or the timed CPU baseline), never as the thing measured or shipped.
 *
 * Data layout is the reference ABI (FK20-CUDA fr.cuh:24, fp.cuh:14, g1.cuh:20-30):
 *   fr  = uint64_t[4]  little-endian limbs, standard (non-Montgomery) form
 *   fp  = uint64_t[6]  little-endian limbs, standard form
 *   g1p = uint64_t[18] x[6], y[6], z[6]  homogeneous projective, infinity = (0,1,0)
 *   g1a = uint64_t[12] x[6], y[6]        affine, infinity = (0,0)
 * Inputs may be any representative (fr: any 256-bit value, fp: any 384-bit value);
 * outputs are canonical (< modulus).
 *
 * Which reference function each routine follows is cited in fk20_oracle.c.
 */
#ifndef FK20_ORACLE_H
#define FK20_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Fr (scalar field, r = 0x73eda753...00000001) ---- */
void or_fr_reduce(uint64_t z[4], const uint64_t x[4]);
void or_fr_add(uint64_t z[4], const uint64_t x[4], const uint64_t y[4]);
void or_fr_sub(uint64_t z[4], const uint64_t x[4], const uint64_t y[4]);
void or_fr_neg(uint64_t z[4], const uint64_t x[4]);
void or_fr_mul(uint64_t z[4], const uint64_t x[4], const uint64_t y[4]);
void or_fr_sqr(uint64_t z[4], const uint64_t x[4]);
void or_fr_inv(uint64_t z[4], const uint64_t x[4]);
void or_fr_pow(uint64_t z[4], const uint64_t x[4], const uint64_t e[4]);
int  or_fr_eq(const uint64_t x[4], const uint64_t y[4]);
/* primitive root of unity of order n (n | 2^32): 5^((r-1)/n) */
void or_fr_root_of_unity(uint64_t z[4], uint64_t n);
/* The reference's 515-entry twiddle table: w^i (i<=512, w of order 512), 2^-9, 2^-9 * w^-128 */
void or_fr_roots515(uint64_t *out);  /* 515 * 4 limbs */

/* ---- Fp (base field, p = 0x1a0111ea...ffffaaab) ---- */
void or_fp_reduce(uint64_t z[6], const uint64_t x[6]);
void or_fp_add(uint64_t z[6], const uint64_t x[6], const uint64_t y[6]);
void or_fp_sub(uint64_t z[6], const uint64_t x[6], const uint64_t y[6]);
void or_fp_neg(uint64_t z[6], const uint64_t x[6]);
void or_fp_mul(uint64_t z[6], const uint64_t x[6], const uint64_t y[6]);
void or_fp_sqr(uint64_t z[6], const uint64_t x[6]);
void or_fp_inv(uint64_t z[6], const uint64_t x[6]);
int  or_fp_eq(const uint64_t x[6], const uint64_t y[6]);

/* ---- G1 ---- */
void or_g1p_inf(uint64_t p[18]);
void or_g1p_gen(uint64_t p[18]);
int  or_g1p_is_inf(const uint64_t p[18]);
int  or_g1p_is_point(const uint64_t p[18]);
int  or_g1p_eq(const uint64_t p[18], const uint64_t q[18]);
void or_g1p_add(uint64_t r[18], const uint64_t p[18], const uint64_t q[18]);
void or_g1p_sub(uint64_t r[18], const uint64_t p[18], const uint64_t q[18]);
void or_g1p_dbl(uint64_t r[18], const uint64_t p[18]);
void or_g1p_neg(uint64_t r[18], const uint64_t p[18]);
void or_g1p_add_affine(uint64_t r[18], const uint64_t p[18], const uint64_t a[12]);
/* k*P by MSB-first double-and-add over the raw 256-bit k */
void or_g1p_mul(uint64_t r[18], const uint64_t p[18], const uint64_t k[4]);
void or_g1p_to_affine(uint64_t a[12], const uint64_t p[18]);
void or_g1p_to_affine_n(uint64_t *a, const uint64_t *p, int64_t count);
/* canonical projective with z in {0,1}: (x,y,1) or (0,1,0) */
void or_g1p_normalize(uint64_t r[18], const uint64_t p[18]);

/* ---- Transforms (natural-order DFT with w = 5^((r-1)/n); inverse includes 1/n) ---- */
void or_fr_fft(uint64_t *out, const uint64_t *in, int n, int inverse);
void or_g1p_fft(uint64_t *out, const uint64_t *in, int n, int inverse);

/* ---- FK20 stages; N = polynomial length (1024/2048/4096), n2 = 2k = N/8 (l = 16) ---- */
void or_fk20_poly2toeplitz_coefficients(uint64_t *tc, const uint64_t *poly, int N);
void or_fk20_msm(uint64_t *hext_fft, const uint64_t *tc_fft, const uint64_t *xext_fft, int n2);
void or_fk20_hext_fft2h_fft(uint64_t *h_fft, const uint64_t *hext_fft, int n2);
/* Whole per-row path for `rows` rows, rows split over `nthreads` threads. */
void or_fk20_poly2h_fft(uint64_t *h_fft, const uint64_t *poly, const uint64_t *xext_fft,
                        int N, int rows, int nthreads);

/* ---- Setup ---- */
/* srs[i] = [tau^i] G for i < count, normalised (z = 1). */
void or_setup_srs(uint64_t *srs, const uint64_t tau[4], int count);
/* commit[rows]: sum_i poly[r][i] * srs[i] (kzg_proofs.py:159 commit_to_poly), i < N. */
void or_commit(uint64_t *commit, const uint64_t *poly, const uint64_t *srs, int N, int rows);
/* xext_fft[16][n2] from the SRS by FK20Py toeplitz_part1 (setup[N-l-1-i::-l] + inf, padded). */
void or_setup2xext_fft(uint64_t *xext_fft, const uint64_t *srs, int N);
/* Same xext_fft computed as [xhat]G from the secret (scalar shadow); normalised points. */
void or_xext_fft_from_tau(uint64_t *xext_fft, const uint64_t tau[4], int N);

/* ---- Closed forms (scalar shadow and KZG openings) ---- */
/* Runs the FK20 algorithm on Fr scalars with setup scalars tau^i: h_fft[k] = [s_fft[k]] G. */
void or_fk20_shadow(uint64_t *s_fft, const uint64_t *poly, const uint64_t tau[4], int N);
/* q(tau) where q = f div (X^16 - c) for the length-N polynomial f. */
void or_kzg_quotient_at(uint64_t out[4], const uint64_t *poly, int N, const uint64_t c[4],
                        const uint64_t tau[4]);

/* ---- Bulk closed forms (threaded; nthreads <= 0: all online CPUs) ---- */
/* aff[i] = canonical affine [s_i] G (infinity (0,0)); s_i any 256-bit value. */
void or_g1_gen_mul_affine_n(uint64_t *aff, const uint64_t *scalars, int64_t count, int nthreads);
/* proj[i] = [s_i] G as the projective representative (z_i X, z_i Y, z_i Z) of the fixed-base
 * sum (X, Y, Z), z_i a non-zero standard-form Fp element; infinity becomes (0, z_i, 0). */
void or_g1_gen_mul_proj_n(uint64_t *proj, const uint64_t *scalars, const uint64_t *zs, int64_t count,
                          int nthreads);
/* Number of i with p[i] (projective, any representative) != aff[i] (affine, (0,0) = infinity);
 * *first receives the smallest such i (or -1). */
int64_t or_g1p_match_affine_n(const uint64_t *p, const uint64_t *aff, int64_t count, int nthreads,
                              int64_t *first);
/* or_fk20_shadow of `rows` polynomials [rows][N] -> s_fft[rows][N/8]. */
void or_fk20_shadow_rows(uint64_t *s_fft, const uint64_t *poly, const uint64_t tau[4], int N, int rows,
                         int nthreads);
/* Scalar shadow of fk20_msm: out[rows][n2] = sum_i tc_fft[rows][i][n2] * xs[i][n2] over Fr
 * (the MSM of bases [xs]G is [out]G). */
void or_fk20_msm_shadow(uint64_t *out, const uint64_t *tc_fft, const uint64_t *xs, int n2, int rows,
                        int nthreads);

/* Number of threads or_fk20_poly2h_fft would use for nthreads <= 0 (online CPUs). */
int or_default_threads(void);

#ifdef __cplusplus
}
#endif
#endif
