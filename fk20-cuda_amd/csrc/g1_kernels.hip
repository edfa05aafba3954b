// G1 kernels of the drop-in entry points besides the transforms (which run on the
// column-batched stage kernels of fast_kernels.hip): the MSM over per-call bases, the
// h_fft upper-half fill, setup gather, affine conversion, compression and element-wise
// G1/Fp arithmetic.
#include <algorithm>

#include "common.hip.h"

namespace fk20 {

__device__ __forceinline__ void words_load(uint32_t *dst, const uint32_t *src, int nwords) {
    const uint4 *s = reinterpret_cast<const uint4 *>(src);
    for (int i = 0; i < nwords / 4; i++) {
        uint4 v = s[i];
        dst[4 * i] = v.x; dst[4 * i + 1] = v.y; dst[4 * i + 2] = v.z; dst[4 * i + 3] = v.w;
    }
}

__device__ __forceinline__ void g1_lds_load(g1p &p, const uint32_t *lds) {
    uint32_t w[36];
#pragma unroll
    for (int i = 0; i < 9; i++) {
        uint4 v = reinterpret_cast<const uint4 *>(lds)[i];
        w[4 * i] = v.x; w[4 * i + 1] = v.y; w[4 * i + 2] = v.z; w[4 * i + 3] = v.w;
    }
    g1_load_packed(p, w);
}

__device__ __forceinline__ void g1_lds_store(uint32_t *lds, const g1p &p) {
    uint32_t w[36];
    g1_store_packed(w, p);
#pragma unroll
    for (int i = 0; i < 9; i++)
        reinterpret_cast<uint4 *>(lds)[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
}

__device__ __forceinline__ void g1_gload_std(g1p &p, const uint32_t *g) {
    uint32_t w[36];
#pragma unroll
    for (int i = 0; i < 9; i++) {
        uint4 v = reinterpret_cast<const uint4 *>(g)[i];
        w[4 * i] = v.x; w[4 * i + 1] = v.y; w[4 * i + 2] = v.z; w[4 * i + 3] = v.w;
    }
    g1_load_std(p, w);
}

__device__ __forceinline__ void g1_gstore_std(uint32_t *g, const g1p &p) {
    uint32_t w[36];
    g1_store_std(w, p);
#pragma unroll
    for (int i = 0; i < 9; i++)
        reinterpret_cast<uint4 *>(g)[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
}

// Scalar multiplication with a non-unrolled bit loop (one inlined dbl and add); any 256-bit k.
__device__ __noinline__ void g1_mul_dev(g1p &r, const g1p &p, const uint32_t *k) {
    g1p q;
    g1_inf(q);
#pragma unroll 1
    for (int w = 7; w >= 0; w--) {
        const uint32_t kw = k[w];
#pragma unroll 1
        for (int b = 31; b >= 0; b--) {
            g1_dbl(q, q);
            if ((kw >> b) & 1) g1_add(q, q, p);
        }
    }
    r = q;
}

// h[row][n/2..n) = (0, 1, 0)  (fk20_poly2h_fft.cu:27)
__global__ void k_hext2h(uint32_t *h, unsigned rows, unsigned n2) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned hh = n2 / 2;
    if (e >= (size_t)rows * hh) return;
    const size_t row = e / hh, j = hh + e % hh;
    uint4 *q = reinterpret_cast<uint4 *>(h + (row * n2 + j) * G1P_WORDS);
#pragma unroll
    for (int i = 0; i < 9; i++) q[i] = make_uint4(0, 0, 0, 0);
    q[3] = make_uint4(1, 0, 0, 0);  // y[0] = 1 (word 12)
}

// Standard-form points -> packed Montgomery points (internal operand form).
__global__ void k_std_to_packed(uint32_t *packed, const uint32_t *pts, size_t count) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    g1p p;
    g1_gload_std(p, pts + e * G1P_WORDS);
    g1_lds_store(packed + e * 36, p);
}

// hext_fft[row][j] = sum_i tc_fft[row][i][j] * xext[i][j] (fk20_msm.cu:23-66), for any 256-bit
// scalars, in two kernels.
//
// k_msm_digits rewrites every scalar as 43 signed radix-64 Booth digits
//     d_w = b[6w-1] + b[6w] + 2 b[6w+1] + 4 b[6w+2] + 8 b[6w+3] + 16 b[6w+4] - 32 b[6w+5]
// (b[-1] = 0, b[256..] = 0), d_w in [-32, 32] and sum_w d_w 64^w = k: each digit reads 7 bits
// and no carries.  Digits go to digits[j][i][w][row] (int8), so the MSM's 64 lanes (rows)
// read 64 consecutive bytes per (term, window).
//
// k_msm_booth: a workgroup is one column j for up to 256 rows (one row per lane), so the
// column's 16 bases are shared: their multiples 1..32 X_i are built once into LDS (16 x 32
// points, 72 KiB) and every lane runs Straus over its 16 digit strings -- 42 windows of 6
// doublings and 43 x 16 complete additions in all, the same steps in every lane (digit 0
// adds infinity).
static constexpr unsigned MSM_NWIN = 43, MSM_HALF = 32;

// Scalar (row, i, j) at element row * row_stride + i * i_stride + j * j_stride; to_D: the
// output is D[bitrev(j)][row] in packed Montgomery form, else hext_fft[row][j] standard form.
struct MsmLayout {
    unsigned row_stride, i_stride, j_stride, to_D;
};

__global__ void __launch_bounds__(256) k_msm_digits(int8_t *digits, const uint32_t *tc, unsigned n2, unsigned rows,
                                                    MsmLayout m) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;  // (j, i, row), row fastest
    if (e >= (size_t)n2 * 16 * rows) return;
    const unsigned row = (unsigned)(e % rows), i = (unsigned)((e / rows) % 16), j = (unsigned)(e / rows / 16);
    uint32_t k[9];
    words_load(k, tc + ((size_t)row * m.row_stride + (size_t)i * m.i_stride + (size_t)j * m.j_stride) * FR_WORDS, 8);
    // K = 2k (257 bits, k[8] = the carry out), so field w is bits 6w .. 6w+6 of K
    k[8] = k[7] >> 31;
#pragma unroll
    for (int q = 7; q > 0; q--) k[q] = (k[q] << 1) | (k[q - 1] >> 31);
    k[0] <<= 1;
    int8_t *out = digits + ((size_t)j * 16 + i) * MSM_NWIN * rows + row;
#pragma unroll
    for (unsigned w = 0; w < MSM_NWIN; w++) {
        const unsigned p = 6 * w, q = p / 32, sh = p % 32;
        uint32_t f = k[q] >> sh;
        if (sh > 25 && q < 8) f |= k[q + 1] << (32 - sh);
        f &= 127u;
        const int d = (int)(f & 1u) + (int)((f >> 1) & 31u) - (int)((f >> 6) & 1u) * 32;
        out[(size_t)w * rows] = (int8_t)d;
    }
}

__device__ __forceinline__ void msm_booth_add(g1p &acc, const uint32_t *tab, int dg) {
    const unsigned a = (unsigned)(dg < 0 ? -dg : dg);
    g1p X;
    g1_lds_load(X, tab + (a ? a - 1 : 0) * 36);
    if (dg < 0) fe_neg<2>(X.y, X.y);
    if (a == 0) g1_inf(X);
    g1_add(acc, acc, X);
}

__global__ void __launch_bounds__(256, 2) k_msm_booth(uint32_t *out, const int8_t *digits, const uint32_t *xpk,
                                                     unsigned n2, unsigned rows, unsigned to_D) {
    __shared__ __attribute__((aligned(16))) uint32_t lds[16 * MSM_HALF * 36];
    const unsigned j = blockIdx.x, t = threadIdx.x;
    // entry e = 32 i + (m - 1) holds m X_i, by double-and-add over m's 6 bits
    for (unsigned e = t; e < 16 * MSM_HALF; e += blockDim.x) {
        const unsigned i = e / MSM_HALF, mult = e % MSM_HALF + 1;
        g1p X, acc;
        g1_lds_load(X, xpk + ((size_t)i * n2 + j) * 36);  // packed Montgomery points in HBM
        g1_inf(acc);
#pragma unroll 1
        for (int b = 5; b >= 0; b--) {
            g1_dbl(acc, acc);
            if ((mult >> b) & 1) g1_add(acc, acc, X);
        }
        g1_lds_store(lds + e * 36, acc);
    }
    __syncthreads();
    const unsigned row0 = blockIdx.y * blockDim.x + t;
    const unsigned row = min(row0, rows - 1);  // lanes past the last row repeat it
    const int8_t *dg = digits + (size_t)j * 16 * MSM_NWIN * rows + row;
    g1p acc;
    g1_inf(acc);
#pragma unroll 1
    for (int w = MSM_NWIN - 1; w >= 0; w--) {
        if (w != (int)MSM_NWIN - 1) {
#pragma unroll 1
            for (int d = 0; d < 6; d++) g1_dbl(acc, acc);
        }
#pragma unroll 1
        for (unsigned i = 0; i < 16; i++)
            msm_booth_add(acc, lds + i * MSM_HALF * 36, dg[((size_t)i * MSM_NWIN + w) * rows]);
    }
    if (row0 < rows) {
        if (to_D) g1_lds_store(out + ((size_t)bitrev(j, 31 - __builtin_clz(n2)) * rows + row) * 36, acc);
        else g1_gstore_std(out + ((size_t)row * n2 + j) * G1P_WORDS, acc);
    }
}

// The same MSM for a few rows (at most MSM_SMALL_ROWS), where latency counts: a lane per
// (term, row), 16 terms x 4 rows per wave, so each lane runs one term's ladder (42 windows
// of 6 doublings, 43 additions) instead of all 16 terms', and the 16 partial sums of a row
// are added in LDS (4 levels).  Lane (i, r) builds the multiples (r+1) X_i + 4k X_i of its
// term's base into LDS first.
static constexpr unsigned MSM_SMALL_ROWS = 16;

__global__ void __launch_bounds__(64) k_msm_booth_small(uint32_t *out, const int8_t *digits, const uint32_t *xpk,
                                                        unsigned n2, unsigned rows, unsigned to_D) {
    __shared__ __attribute__((aligned(16))) uint32_t lds[16 * MSM_HALF * 36];
    const unsigned j = blockIdx.x, t = threadIdx.x, i = t % 16, r = t / 16;
    {
        g1p X, step, e;
        g1_lds_load(X, xpk + ((size_t)i * n2 + j) * 36);  // packed Montgomery points in HBM
        g1_dbl(step, X);                                  // 2 X
        if (r == 0) e = X;
        else if (r == 1) e = step;
        else if (r == 2) g1_add(e, step, X);
        else g1_dbl(e, step);
        g1_dbl(step, step);                               // 4 X
#pragma unroll 1
        for (unsigned k = 0; k < MSM_HALF / 4; k++) {     // entry m - 1 holds m X_i, m = r + 1 + 4k
            g1_lds_store(lds + (i * MSM_HALF + r + 4 * k) * 36, e);
            if (k + 1 < MSM_HALF / 4) g1_add(e, e, step);
        }
    }
    __syncthreads();
    const unsigned row0 = blockIdx.y * 4 + r;
    const unsigned row = min(row0, rows - 1);  // lanes past the last row repeat it
    const int8_t *dg = digits + ((size_t)j * 16 + i) * MSM_NWIN * rows + row;
    g1p acc;
    g1_inf(acc);
#pragma unroll 1
    for (int w = MSM_NWIN - 1; w >= 0; w--) {
        if (w != (int)MSM_NWIN - 1) {
#pragma unroll 1
            for (int d = 0; d < 6; d++) g1_dbl(acc, acc);
        }
        msm_booth_add(acc, lds + i * MSM_HALF * 36, dg[(size_t)w * rows]);
    }
    __syncthreads();  // the table is dead: its first 64 slots take the partial sums
    g1_lds_store(lds + t * 36, acc);
#pragma unroll 1
    for (unsigned h = 8; h >= 1; h /= 2) {
        __syncthreads();
        g1p o;
        if (i < h) {
            g1_lds_load(o, lds + (t + h) * 36);
            g1_add(acc, acc, o);
        }
        __syncthreads();
        if (i < h) g1_lds_store(lds + t * 36, acc);
    }
    if (i == 0 && row0 < rows) {
        if (to_D) g1_lds_store(out + ((size_t)bitrev(j, 31 - __builtin_clz(n2)) * rows + row) * 36, acc);
        else g1_gstore_std(out + ((size_t)row * n2 + j) * G1P_WORDS, acc);
    }
}

// Issue-rate probe of the 32 x 32 -> 64-bit multiply-add every Fp product is built from: 8
// independent accumulator chains per lane, each step acc_k = x_k * lo(acc_(k+1)) + acc_k,
// so every instruction is one v_mad_u64_u32 whose operands change each iteration.
__global__ void __launch_bounds__(256) k_mad_probe(uint64_t *out, unsigned iters) {
    uint64_t acc[8];
    uint32_t x[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        acc[k] = (uint64_t)(threadIdx.x + 977u * k) << 7 | k;
        x[k] = 0x9e3779b9u * (blockIdx.x + 1) + 0x85ebca6bu * k;
    }
#pragma unroll 1
    for (unsigned it = 0; it < iters; it++) {
#pragma unroll
        for (int k = 0; k < 8; k++) acc[k] = (uint64_t)x[k] * (uint32_t)acc[(k + 1) & 7] + acc[k];
    }
    uint64_t h = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) h ^= acc[k];
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = h;
}

// ---------------------------------------------------------------- limb-layout probe
//
// The Fp multiplication in the two limb layouts considered for the library: LIMBS29 = the
// library's own 14 x 29-bit product scanning (fp.hip.h fe_mul, Montgomery radix 2^406, one
// v_mad_u64_u32 per partial product into a carry-free 64-bit column), LIMBS32 = 12 x 32-bit
// CIOS (Koc, Acar, Kaliski 1996; radix 2^384), where each partial product needs a 64-bit
// multiply-add plus a carry addition.  Both lazy: inputs and outputs below 2p.

// z = a b 2^-384 mod p, 12 x 32-bit CIOS; a, b < 2p (4p < 2^384, so no final subtraction).
__device__ __forceinline__ void fp32_mul_cios(uint32_t *z, const uint32_t *a, const uint32_t *b) {
    uint32_t t[14];
#pragma unroll
    for (int j = 0; j < 14; j++) t[j] = 0;
#pragma unroll
    for (int i = 0; i < 12; i++) {
        uint64_t c = 0;
#pragma unroll
        for (int j = 0; j < 12; j++) {
            const uint64_t s = (uint64_t)a[j] * b[i] + t[j] + c;
            t[j] = (uint32_t)s;
            c = s >> 32;
        }
        uint64_t s = (uint64_t)t[12] + c;
        t[12] = (uint32_t)s;
        t[13] = (uint32_t)(s >> 32);
        const uint32_t m = t[0] * FP32_PINV;
        s = (uint64_t)m * FP32_P[0] + t[0];
        c = s >> 32;
#pragma unroll
        for (int j = 1; j < 12; j++) {
            s = (uint64_t)m * FP32_P[j] + t[j] + c;
            t[j - 1] = (uint32_t)s;
            c = s >> 32;
        }
        s = (uint64_t)t[12] + c;
        t[11] = (uint32_t)s;
        t[12] = t[13] + (uint32_t)(s >> 32);
    }
#pragma unroll
    for (int j = 0; j < 12; j++) z[j] = t[j];
}

// Lane l takes x = xs[l mod 1024] and the shared y, runs two independent chains of `iters`
// multiplications by y, and writes their product (12 words, packed).
template <int LIMBS>
__global__ void __launch_bounds__(256) k_fp_mul_probe(uint32_t *out, const uint32_t *xs, const uint32_t *y,
                                                      unsigned iters) {
    const size_t lane = (size_t)blockIdx.x * 256 + threadIdx.x;
    const uint32_t *x = xs + (lane & 1023) * 12;
    if (LIMBS == 29) {
        fe a, b, c;
        fe_unpack(a, x);
        fe_unpack(b, y);
        c = a;
        fe_mul(c, c, b);  // a second chain, offset by one step
#pragma unroll 1
        for (unsigned it = 0; it < iters; it++) {
            fe_mul(a, a, b);
            fe_mul(c, c, b);
        }
        fe_mul(a, a, c);
        fe_pack(out + lane * 12, a);
    } else {
        uint32_t a[12], b[12], c[12];
#pragma unroll
        for (int j = 0; j < 12; j++) {
            a[j] = x[j];
            b[j] = y[j];
        }
        fp32_mul_cios(c, a, b);
#pragma unroll 1
        for (unsigned it = 0; it < iters; it++) {
            fp32_mul_cios(a, a, b);
            fp32_mul_cios(c, c, b);
        }
        fp32_mul_cios(a, a, c);
#pragma unroll
        for (int j = 0; j < 12; j++) out[lane * 12 + j] = a[j];
    }
}

// xext[i][m] = srs[N-17-i-16m] for m <= k-2, infinity otherwise (FK20Py fk20_multi.py:82-85:
// setup[n-l-1-i::-l] + [Z1], padded to 2k with Z1).
__global__ void k_setup_gather(uint32_t *xext, const uint32_t *srs, unsigned N) {
    const unsigned n2 = N / 8, k = N / 16;
    const unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 16 * n2) return;
    const unsigned i = e / n2, m = e % n2;
    uint4 *q = reinterpret_cast<uint4 *>(xext + (size_t)e * G1P_WORDS);
    if (m + 2 <= k) {
        const uint4 *src = reinterpret_cast<const uint4 *>(srs + (size_t)(N - 17 - i - 16 * m) * G1P_WORDS);
#pragma unroll
        for (int w = 0; w < 9; w++) q[w] = src[w];
    } else {
#pragma unroll
        for (int w = 0; w < 9; w++) q[w] = make_uint4(0, 0, 0, 0);
        q[3] = make_uint4(1, 0, 0, 0);
    }
}

// Canonical affine (x, y); infinity -> (0, 0) as g1a_fromG1p (g1a.cu:48-56) gives.
__global__ void k_to_affine(uint32_t *out, const uint32_t *in, size_t count) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    g1p p;
    g1_gload_std(p, in + e * G1P_WORDS);
    fe x, y;
    g1_to_affine(x, y, p);
    fe_from_mont(x, x);
    fe_from_mont(y, y);
    uint32_t w[24];
    fe_pack(w, x);
    fe_pack(w + 12, y);
    uint4 *q = reinterpret_cast<uint4 *>(out + e * G1A_WORDS);
#pragma unroll
    for (int i = 0; i < 6; i++) q[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
}

__global__ void k_compress(uint32_t *out, const uint32_t *in, size_t count) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    g1p p;
    g1_gload_std(p, in + e * G1P_WORDS);
    uint32_t w[12];
    g1_compress(w, p);
    uint4 *q = reinterpret_cast<uint4 *>(out + e * 12);
#pragma unroll
    for (int i = 0; i < 3; i++) q[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
}

// ---------------------------------------------------------------- element-wise ops

__device__ __forceinline__ void fe_gload_std(fe &a, const uint32_t *g) {
    uint32_t w[12];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        uint4 v = reinterpret_cast<const uint4 *>(g)[i];
        w[4 * i] = v.x; w[4 * i + 1] = v.y; w[4 * i + 2] = v.z; w[4 * i + 3] = v.w;
    }
    fe t;
    fe_unpack(t, w);
    fe_to_mont(a, t);
}

__device__ __forceinline__ void fe_gstore_std(uint32_t *g, const fe &a) {
    fe t;
    fe_from_mont(t, a);
    uint32_t w[12];
    fe_pack(w, t);
#pragma unroll
    for (int i = 0; i < 3; i++)
        reinterpret_cast<uint4 *>(g)[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
}

__global__ void k_fp_op(int op, uint32_t *out, const uint32_t *a, const uint32_t *b, size_t count) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    fe x, y, z;
    fe_gload_std(x, a + e * 12);
    fe_zero(y);
    if (op <= 2 && b) fe_gload_std(y, b + e * 12);  // binary ops only
    switch (op) {
    case 0: fe_add(z, x, y); break;
    case 1: fe_sub<2>(z, x, y); break;
    case 2: fe_mul(z, x, y); break;
    case 3: fe_sqr(z, x); break;
    case 4: fe_neg<2>(z, x); break;
    case 5: fe_inv(z, x); break;
    default: z = x; break;
    }
    fe_gstore_std(out + e * 12, z);
}

__global__ void k_g1_op(int op, uint32_t *out, const uint32_t *pp, const uint32_t *qq, const uint32_t *kk,
                        size_t count) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    g1p p, q, r;
    g1_gload_std(p, pp + e * G1P_WORDS);
    switch (op) {
    case 0: g1_gload_std(q, qq + e * G1P_WORDS); g1_add(r, p, q); break;
    case 1: g1_gload_std(q, qq + e * G1P_WORDS); g1_neg(q, q); g1_add(r, p, q); break;
    case 2: g1_dbl(r, p); break;
    case 3: {
        uint32_t k[8];
#pragma unroll
        for (int i = 0; i < 8; i++) k[i] = kk[e * FR_WORDS + i];
        g1_mul_dev(r, p, k);
    } break;
    case 4: g1_neg(r, p); break;
    case 5: {
        fe ax, ay;
        fe_gload_std(ax, qq + e * G1A_WORDS);
        fe_gload_std(ay, qq + e * G1A_WORDS + 12);
        fe cx, cy;
        fe_canonical(cx, ax);
        fe_canonical(cy, ay);
        if (fe_is_zero_canonical(cx) && fe_is_zero_canonical(cy)) r = p;
        else g1_add_aff(r, p, ax, ay);
    } break;
    case 6:
    case 7: {
        // 6: y^2 z == x^3 + 4 z^3 (or infinity); 7: cross-multiplied equality
        bool res;
        if (op == 6) {
            fe l, rr, t, zc, xc, yc;
            fe_canonical(zc, p.z);
            fe_canonical(xc, p.x);
            fe_canonical(yc, p.y);
            if (fe_is_zero_canonical(zc)) {
                res = fe_is_zero_canonical(xc) && !fe_is_zero_canonical(yc);
            } else {
                fe_sqr(l, p.y); fe_mul(l, l, p.z);
                fe_sqr(rr, p.x); fe_mul(rr, rr, p.x);
                fe_sqr(t, p.z); fe_mul(t, t, p.z); fe_mul_small<4>(t, t);
                fe_add(rr, rr, t);
                fe_sub<16>(l, l, rr);  // rr < 10p
                fe_canonical(l, l);
                res = fe_is_zero_canonical(l);
            }
        } else {
            g1_gload_std(q, qq + e * G1P_WORDS);
            fe za, zb, a1, b1;
            fe_canonical(za, p.z);
            fe_canonical(zb, q.z);
            const bool ia = fe_is_zero_canonical(za), ib = fe_is_zero_canonical(zb);
            if (ia || ib) {
                res = ia && ib;
            } else {
                fe_mul(a1, p.x, q.z); fe_mul(b1, q.x, p.z);
                fe_sub<2>(a1, a1, b1); fe_canonical(a1, a1);
                res = fe_is_zero_canonical(a1);
                fe_mul(a1, p.y, q.z); fe_mul(b1, q.y, p.z);
                fe_sub<2>(a1, a1, b1); fe_canonical(a1, a1);
                res = res && fe_is_zero_canonical(a1);
            }
        }
        uint4 *o = reinterpret_cast<uint4 *>(out + e * G1P_WORDS);
#pragma unroll
        for (int i = 0; i < 9; i++) o[i] = make_uint4(0, 0, 0, 0);
        o[0].x = res ? 1u : 0u;
        return;
    }
    default: r = p; break;
    }
    g1_gstore_std(out + e * G1P_WORDS, r);
}

// ---------------------------------------------------------------- launchers

hipError_t launch_hext2h(uint32_t *h, unsigned rows, unsigned n2, hipStream_t s) {
    const size_t total = (size_t)rows * (n2 / 2);
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_hext2h, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, h, rows, n2);
    return hipGetLastError();
}

hipError_t launch_std_to_packed(uint32_t *packed, const uint32_t *pts, size_t count, hipStream_t s) {
    if (!count) return hipSuccess;
    hipLaunchKernelGGL(k_std_to_packed, dim3((unsigned)((count + 63) / 64)), dim3(64), 0, s, packed, pts, count);
    return hipGetLastError();
}

size_t msm_plain_digit_bytes(unsigned rows, unsigned n2) { return (size_t)rows * n2 * 16 * MSM_NWIN; }

static hipError_t msm_plain(uint32_t *out, const uint32_t *tc, const uint32_t *xpk, int8_t *digits, unsigned rows,
                            unsigned n2, MsmLayout m, hipStream_t s) {
    if (!rows) return hipSuccess;
    const size_t scalars = (size_t)rows * n2 * 16;
    hipLaunchKernelGGL(k_msm_digits, dim3((unsigned)((scalars + 255) / 256)), dim3(256), 0, s, digits, tc, n2, rows, m);
    if (rows <= MSM_SMALL_ROWS) {
        hipLaunchKernelGGL(k_msm_booth_small, dim3(n2, (rows + 3) / 4), dim3(64), 0, s, out, (const int8_t *)digits,
                           xpk, n2, rows, m.to_D);
        return hipGetLastError();
    }
    const unsigned threads = 64 * std::min(4u, (rows + 63) / 64);
    hipLaunchKernelGGL(k_msm_booth, dim3(n2, (rows + threads - 1) / threads), dim3(threads), 0, s, out,
                       (const int8_t *)digits, xpk, n2, rows, m.to_D);
    return hipGetLastError();
}

hipError_t launch_msm_plain(uint32_t *hext, const uint32_t *tc_fft, const uint32_t *xpk, int8_t *digits,
                            unsigned rows, unsigned n2, hipStream_t s) {
    // scalars tc_fft[row][i][j], output hext_fft[row][j] (standard form)
    return msm_plain(hext, tc_fft, xpk, digits, rows, n2, MsmLayout{16u * n2, n2, 1u, 0u}, s);
}

hipError_t launch_msm_plain_T(uint32_t *D, const uint32_t *tcT, const uint32_t *xpk, int8_t *digits, unsigned rows,
                              unsigned n2, hipStream_t s) {
    // scalars tcT[j][i][row], output D[bitrev(j)][row] (packed Montgomery)
    return msm_plain(D, tcT, xpk, digits, rows, n2, MsmLayout{1u, rows, 16u * rows, 1u}, s);
}

hipError_t launch_mad_probe(uint64_t *out, unsigned blocks, unsigned iters, hipStream_t s) {
    hipLaunchKernelGGL(k_mad_probe, dim3(blocks), dim3(256), 0, s, out, iters);
    return hipGetLastError();
}

hipError_t launch_fp_mul_probe(int limbs, uint32_t *out, const uint32_t *xs, const uint32_t *y, unsigned blocks,
                               unsigned iters, hipStream_t s) {
    if (limbs == 29) hipLaunchKernelGGL(k_fp_mul_probe<29>, dim3(blocks), dim3(256), 0, s, out, xs, y, iters);
    else hipLaunchKernelGGL(k_fp_mul_probe<32>, dim3(blocks), dim3(256), 0, s, out, xs, y, iters);
    return hipGetLastError();
}

hipError_t launch_setup_gather(uint32_t *xext, const uint32_t *srs, unsigned N, hipStream_t s) {
    const unsigned total = 16 * (N / 8);
    hipLaunchKernelGGL(k_setup_gather, dim3((total + 255) / 256), dim3(256), 0, s, xext, srs, N);
    return hipGetLastError();
}

hipError_t launch_to_affine(uint32_t *out, const uint32_t *in, size_t count, hipStream_t s) {
    if (!count) return hipSuccess;
    hipLaunchKernelGGL(k_to_affine, dim3((unsigned)((count + 63) / 64)), dim3(64), 0, s, out, in, count);
    return hipGetLastError();
}

hipError_t launch_compress(uint32_t *out, const uint32_t *in, size_t count, hipStream_t s) {
    if (!count) return hipSuccess;
    hipLaunchKernelGGL(k_compress, dim3((unsigned)((count + 63) / 64)), dim3(64), 0, s, out, in, count);
    return hipGetLastError();
}

hipError_t launch_fp_op(int op, uint32_t *out, const uint32_t *a, const uint32_t *b, size_t count,
                        hipStream_t s) {
    if (!count) return hipSuccess;
    hipLaunchKernelGGL(k_fp_op, dim3((unsigned)((count + 63) / 64)), dim3(64), 0, s, op, out, a, b, count);
    return hipGetLastError();
}

hipError_t launch_g1_op(int op, uint32_t *out, const uint32_t *p, const uint32_t *q, const uint32_t *k,
                        size_t count, hipStream_t s) {
    if (!count) return hipSuccess;
    hipLaunchKernelGGL(k_g1_op, dim3((unsigned)((count + 63) / 64)), dim3(64), 0, s, op, out, p, q, k, count);
    return hipGetLastError();
}

}  // namespace fk20
