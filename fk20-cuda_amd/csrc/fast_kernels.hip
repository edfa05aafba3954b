// Kernels of the batched FK20 path (many rows sharing one resident setup).
//
//  1. k_scale_bases    xext_fft / n2 -> packed Montgomery bases (the IFFT's 1/n2 is folded
//                      into the setup, so the G1 IFFT below runs unscaled).
//  2. k_table_build    fixed-base signed-window tables T[b][win][d] = d * 2^(w*win) * base_b
//                      (d = 1 .. 2^(w-1)), affine, canonical Montgomery, one per base.
//  3. k_toeplitz_fft_T Toeplitz coefficients + NTT per (row, i), written transposed as
//                      tcT[j][i][row] so 64 rows of one column are contiguous.
//  4. k_msm_comb       hext_fft[j] of every row = sum_i tc_fft[i][j] * xext_fft[i][j] by
//                      signed fixed windows: ceil(256/w) mixed additions per term, no
//                      doublings.  Lanes = rows of one column, so a wave gathers from the
//                      table region of one base.  Output goes to D[bitrev(j)][row].
//  5. k_dit_stage / k_middle / k_dif_stage
//                      the G1 IFFT (DIT, bit-reversed in, natural out), the fused last IFFT
//                      stage + zero upper half + first FFT stage, and the G1 FFT (DIF,
//                      natural in, bit-reversed out), column-batched: a wave is one
//                      butterfly for 64 rows, so the twiddle is wave-uniform and runs as a
//                      precomputed GLV + wNAF program (uniform control flow, no divergence).
//  6. k_fft_output     D[pos][row] -> h_fft[row][bitrev(pos)] in the reference ABI.
// Same mathematics as fk20_msm.cu:23-66, g1p_fft.cu:25-330 and fk20_hext_fft2h_fft.cu:27-70;
// the layout and algorithms are this library's own.
#include "common.hip.h"
#include "fast.hip.h"

namespace fk20 {

// ---------------------------------------------------------------- memory helpers

__device__ __forceinline__ void ld_words(uint32_t *w, const uint32_t *g, int n4) {
#pragma unroll
    for (int i = 0; i < n4; i++) {
        uint4 v = reinterpret_cast<const uint4 *>(g)[i];
        w[4 * i] = v.x; w[4 * i + 1] = v.y; w[4 * i + 2] = v.z; w[4 * i + 3] = v.w;
    }
}

__device__ __forceinline__ void st_words(uint32_t *g, const uint32_t *w, int n4) {
#pragma unroll
    for (int i = 0; i < n4; i++)
        reinterpret_cast<uint4 *>(g)[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
}

__device__ __forceinline__ void pt_load(g1p &p, const uint32_t *g) {
    uint32_t w[36];
    ld_words(w, g, 9);
    g1_load_packed(p, w);
}

__device__ __forceinline__ void pt_store(uint32_t *g, const g1p &p) {
    uint32_t w[36];
    g1_store_packed(w, p);
    st_words(g, w, 9);
}

__device__ __forceinline__ void fe_load(fe &a, const uint32_t *g) {
    uint32_t w[12];
    ld_words(w, g, 3);
    fe_unpack(a, w);
}

__device__ __forceinline__ void fe_store(uint32_t *g, const fe &a) {
    uint32_t w[12];
    fe_pack(w, a);
    st_words(g, w, 3);
}

// ---------------------------------------------------------------- setup: scaled bases

__device__ __noinline__ void g1_mul_std(g1p &r, const g1p &p, const uint32_t *k) {
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

// xs[b] = scale * xext[b] (packed Montgomery); inf[b] = 1 if xext[b] is infinity.
__global__ void __launch_bounds__(64) k_scale_bases(uint32_t *xs, uint32_t *inf, const uint32_t *xext,
                                                    size_t count, const uint32_t *scale) {
    const size_t b = (size_t)blockIdx.x * 64 + threadIdx.x;
    if (b >= count) return;
    uint32_t w[36];
    ld_words(w, xext + b * G1P_WORDS, 9);
    g1p p;
    g1_load_std(p, w);
    fe zc;
    fe_canonical(zc, p.z);
    inf[b] = fe_is_zero_canonical(zc) ? 1u : 0u;
    uint32_t k[8];
#pragma unroll
    for (int i = 0; i < 8; i++) k[i] = scale[i];
    g1_mul_std(p, p, k);
    pt_store(xs + b * 36, p);
}

// ---------------------------------------------------------------- setup: window tables

// One thread per (base, window) of the chunk [base0, base0 + nb).  Entries are built as
// projective multiples d*B (d = 1..H) and converted to affine with one inversion per thread
// (Montgomery's batch trick).  zt holds (z_d, prefix_{d-1}) per entry for the backward pass.
__global__ void __launch_bounds__(64) k_table_build(uint32_t *table, uint32_t *zt, const uint32_t *xs,
                                                    const uint32_t *inf, unsigned base0, unsigned nb,
                                                    unsigned win_bits, unsigned nwin) {
    const unsigned t = blockIdx.x * 64 + threadIdx.x;
    if (t >= nb * nwin) return;
    const unsigned b = base0 + t / nwin, win = t % nwin;
    const unsigned H = 1u << (win_bits - 1);
    uint32_t *tab = table + ((size_t)b * nwin + win) * H * G1A_WORDS;
    uint32_t *z = zt + (size_t)t * H * 24;
    if (inf[b]) {
        for (unsigned e = 0; e < H; e++) {
            uint32_t zero[24] = {0};
            st_words(tab + (size_t)e * G1A_WORDS, zero, 6);
        }
        return;
    }
    g1p B, E;
    pt_load(B, xs + (size_t)b * 36);
#pragma unroll 1
    for (unsigned d = 0; d < win_bits * win; d++) g1_dbl(B, B);
    E = B;
    fe prefix;
    fe_one(prefix);
#pragma unroll 1
    for (unsigned e = 0; e < H; e++) {
        if (e) g1_add(E, E, B);
        fe_store(tab + (size_t)e * G1A_WORDS, E.x);
        fe_store(tab + (size_t)e * G1A_WORDS + 12, E.y);
        fe_store(z + (size_t)e * 24, E.z);
        fe_store(z + (size_t)e * 24 + 12, prefix);  // prefix of z_0 .. z_{e-1}
        fe_mul(prefix, prefix, E.z);
    }
    fe inv;
    fe_canonical(prefix, prefix);
    fe_inv(inv, prefix);
#pragma unroll 1
    for (int e = (int)H - 1; e >= 0; e--) {
        fe ze, pe, zinv, x, y;
        fe_load(ze, z + (size_t)e * 24);
        fe_load(pe, z + (size_t)e * 24 + 12);
        fe_mul(zinv, inv, pe);
        fe_mul(inv, inv, ze);
        fe_load(x, tab + (size_t)e * G1A_WORDS);
        fe_load(y, tab + (size_t)e * G1A_WORDS + 12);
        fe_mul(x, x, zinv);
        fe_mul(y, y, zinv);
        fe_canonical(x, x);
        fe_canonical(y, y);
        fe_store(tab + (size_t)e * G1A_WORDS, x);
        fe_store(tab + (size_t)e * G1A_WORDS + 12, y);
    }
}

// ---------------------------------------------------------------- Toeplitz + NTT (transposed)

__device__ __forceinline__ int toeplitz_src_T(unsigned i, unsigned j, unsigned N) {
    const unsigned k = N / 16;
    if (j == 0) return (int)(N - 1 - i);
    if (j <= k + 1) return -1;
    return (int)(31 - i + 16 * (j - (k + 2)));
}

template <unsigned LOGN>
__global__ void __launch_bounds__(256) k_toeplitz_fft_T(uint32_t *tcT, const uint32_t *poly, unsigned rows,
                                                        unsigned N, const DevTables *T) {
    constexpr unsigned n = 1u << LOGN, half = n / 2;
    __shared__ fs lds[n];
    const unsigned t = threadIdx.x;
    const unsigned i = blockIdx.x % 16;
    const size_t row = blockIdx.x / 16;
    poly += row * N * FR_WORDS;
    for (unsigned e = t; e < n; e += half) {
        const int src = toeplitz_src_T(i, e, N);
        fs a;
        fs_zero(a);
        if (src >= 0) {
            uint32_t w[8];
            ld_words(w, poly + (size_t)src * FR_WORDS, 2);
#pragma unroll
            for (int q = 0; q < 8; q++) a.l[q] = w[q];
            fs_to_mont(a, a);
        }
        lds[bitrev(e, LOGN)] = a;
    }
    __syncthreads();
    for (unsigned s = 0; s < LOGN; s++) {
        const unsigned h = 1u << s, j = t & (h - 1);
        const unsigned l = ((t >> s) << (s + 1)) + j, r = l + h;
        fs x = lds[r], y = lds[l], w;
        fs_set(w, T->tw_mont[j << (8 - s)]);
        fs_mul(x, x, w);
        fs_add(lds[l], y, x);
        fs_sub(lds[r], y, x);
        __syncthreads();
    }
    for (unsigned e = t; e < n; e += half) {
        fs a = lds[e];
        fs_from_mont(a, a);
        st_words(tcT + (((size_t)e * 16 + i) * rows + row) * FR_WORDS, a.l, 2);
    }
}

// ---------------------------------------------------------------- comb MSM

template <unsigned WIN>
__global__ void __launch_bounds__(64) k_msm_comb(uint32_t *D, const uint32_t *tcT, const uint32_t *table,
                                                 const uint32_t *inf, unsigned rows, unsigned n2,
                                                 unsigned logn2) {
    constexpr unsigned NWIN = (256 + WIN - 1) / WIN, H = 1u << (WIN - 1), MASK = (1u << WIN) - 1;
    const unsigned rg = (rows + 63) / 64;
    const unsigned j = blockIdx.x / rg;
    const unsigned row = (blockIdx.x % rg) * 64 + threadIdx.x;
    if (row >= rows) return;
    g1p acc;
    g1_inf(acc);
#pragma unroll 1
    for (unsigned i = 0; i < 16; i++) {
        const unsigned b = i * n2 + j;
        if (inf[b]) continue;
        uint32_t k[8];
        ld_words(k, tcT + (((size_t)j * 16 + i) * rows + row) * FR_WORDS, 2);
        const uint32_t *tb = table + (size_t)b * NWIN * H * G1A_WORDS;
        uint32_t carry = 0;
#pragma unroll 1
        for (unsigned win = 0; win < NWIN; win++) {
            const uint32_t bits = (k[0] & MASK) + carry;
            carry = bits > H ? 1u : 0u;
            const int d = (int)bits - (int)(carry << WIN);
            // shift the scalar right by WIN bits
#pragma unroll
            for (int q = 0; q < 7; q++) k[q] = (k[q] >> WIN) | (k[q + 1] << (32 - WIN));
            k[7] >>= WIN;
            if (d != 0) {
                const unsigned e = (unsigned)(d < 0 ? -d : d) - 1;
                const uint32_t *ent = tb + ((size_t)win * H + e) * G1A_WORDS;
                fe x, y;
                fe_load(x, ent);
                fe_load(y, ent + 12);
                if (d < 0) fe_neg<2>(y, y);
                g1_add_aff(acc, acc, x, y);
            }
        }
    }
    pt_store(D + ((size_t)bitrev(j, logn2) * rows + row) * 36, acc);
}

// ---------------------------------------------------------------- twiddle programs

// x <- k * x for the wave-uniform scalar whose GLV + wNAF program starts at ops[0]
// (op layout in fast.hip.h).  Table of odd multiples P, 3P, .., (2^(W-1)-1) P in registers.
template <int W>
__device__ __forceinline__ void tw_mul(g1p &x, const uint32_t *ops) {
    constexpr int NT = 1 << (W - 2);
    g1p T0 = x, T1, T2, T3;
    if (NT > 1) {
        g1p x2;
        g1_dbl(x2, x);
        g1_add(T1, T0, x2);
        if (NT > 2) {
            g1_add(T2, T1, x2);
            g1_add(T3, T2, x2);
        }
    }
    fe beta;
    fe_set(beta, FP_BETA);
    g1p acc;
    bool started = false;
#pragma unroll 1
    for (unsigned k = 0;; k++) {
        const uint32_t op = __builtin_amdgcn_readfirstlane(ops[k]);  // wave-uniform program
        const unsigned nd = op & TWOP_DBL_MASK;
#pragma unroll 1
        for (unsigned d = 0; d < nd; d++) g1_dbl(acc, acc);
        if (op & TWOP_END) break;
        const unsigned e = (op >> TWOP_ENTRY_SHIFT) & 7;
        g1p q;
        switch (e) {
        case 0: q = T0; break;
        case 1: q = T1; break;
        case 2: q = T2; break;
        default: q = T3; break;
        }
        if (op & TWOP_PHI) fe_mul(q.x, q.x, beta);
        if (op & TWOP_NEG) fe_neg<16>(q.y, q.y);
        if (started) {
            g1_add(acc, acc, q);
        } else {
            acc = q;
            started = true;
        }
    }
    x = acc;
}

// IFFT decimation-in-time stage s over D[pos][row] (bit-reversed input), twiddles w^-idx.
template <int W>
__global__ void __launch_bounds__(64) k_dit_stage(uint32_t *D, unsigned rows, unsigned s,
                                                  const uint32_t *ops, const uint32_t *index) {
    const unsigned rg = (rows + 63) / 64;
    const unsigned b = blockIdx.x / rg;
    const unsigned row = (blockIdx.x % rg) * 64 + threadIdx.x;
    if (row >= rows) return;
    const unsigned j = b & ((1u << s) - 1);
    const unsigned l = ((b >> s) << (s + 1)) + j, r = l + (1u << s);
    const unsigned idx = j << (8 - s);
    g1p x, y, a;
    pt_load(x, D + ((size_t)r * rows + row) * 36);
    if (idx) tw_mul<W>(x, ops + index[TW_INV * 512 + idx]);
    pt_load(y, D + ((size_t)l * rows + row) * 36);
    g1_add(a, y, x);
    g1_neg(x, x);
    g1_add(x, y, x);
    pt_store(D + ((size_t)l * rows + row) * 36, a);
    pt_store(D + ((size_t)r * rows + row) * 36, x);
}

// Last IFFT stage, upper half dropped (h = first half), then the first FFT DIF stage whose
// upper input half is zero:  x = E_j + w^-idx O_j;  D[j] = x;  D[j + n2/2] = w^idx x.
template <int W>
__global__ void __launch_bounds__(64) k_middle(uint32_t *D, unsigned rows, unsigned logn2,
                                               const uint32_t *ops, const uint32_t *index) {
    const unsigned rg = (rows + 63) / 64;
    const unsigned j = blockIdx.x / rg;
    const unsigned row = (blockIdx.x % rg) * 64 + threadIdx.x;
    if (row >= rows) return;
    const unsigned hn = 1u << (logn2 - 1);
    const unsigned idx = j << (9 - logn2);
    g1p o, e;
    pt_load(o, D + ((size_t)(j + hn) * rows + row) * 36);
    if (idx) tw_mul<W>(o, ops + index[TW_INV * 512 + idx]);
    pt_load(e, D + ((size_t)j * rows + row) * 36);
    g1_add(e, e, o);
    pt_store(D + ((size_t)j * rows + row) * 36, e);
    if (idx) tw_mul<W>(e, ops + index[TW_FWD * 512 + idx]);
    pt_store(D + ((size_t)(j + hn) * rows + row) * 36, e);
}

// FFT decimation-in-frequency stage s >= 1 (span n2 >> (s+1)), twiddles w^idx after the
// subtraction.
template <int W>
__global__ void __launch_bounds__(64) k_dif_stage(uint32_t *D, unsigned rows, unsigned logn2, unsigned s,
                                                  const uint32_t *ops, const uint32_t *index) {
    const unsigned rg = (rows + 63) / 64;
    const unsigned b = blockIdx.x / rg;
    const unsigned row = (blockIdx.x % rg) * 64 + threadIdx.x;
    if (row >= rows) return;
    const unsigned lh = logn2 - s - 1;  // log2 of the span
    const unsigned h = 1u << lh;
    const unsigned j = b & (h - 1);
    const unsigned l = ((b >> lh) << (lh + 1)) + j, r = l + h;
    const unsigned idx = j << (8 - lh);
    g1p u, v, a;
    pt_load(u, D + ((size_t)l * rows + row) * 36);
    pt_load(v, D + ((size_t)r * rows + row) * 36);
    g1_add(a, u, v);
    pt_store(D + ((size_t)l * rows + row) * 36, a);  // stored first: not live across tw_mul
    g1_neg(v, v);
    g1_add(v, u, v);
    if (idx) tw_mul<W>(v, ops + index[TW_FWD * 512 + idx]);
    pt_store(D + ((size_t)r * rows + row) * 36, v);
}

// D[pos][row] (bit-reversed) -> h_fft[row][bitrev(pos)], canonical standard projective.
__global__ void __launch_bounds__(64) k_fft_output(uint32_t *h_fft, const uint32_t *D, unsigned rows,
                                                   unsigned logn2) {
    const unsigned rg = (rows + 63) / 64;
    const unsigned pos = blockIdx.x / rg;
    const unsigned row = (blockIdx.x % rg) * 64 + threadIdx.x;
    if (row >= rows) return;
    g1p p;
    pt_load(p, D + ((size_t)pos * rows + row) * 36);
    uint32_t w[36];
    g1_store_std(w, p);
    st_words(h_fft + ((size_t)row * (1u << logn2) + bitrev(pos, logn2)) * G1P_WORDS, w, 9);
}

// ---------------------------------------------------------------- launchers

static inline unsigned blocks_for(unsigned per_row_units, unsigned rows) {
    return per_row_units * ((rows + 63) / 64);
}

hipError_t launch_scale_bases(uint32_t *xs, uint32_t *inf, const uint32_t *xext, size_t count,
                              const uint32_t *scale, hipStream_t s) {
    hipLaunchKernelGGL(k_scale_bases, dim3((unsigned)((count + 63) / 64)), dim3(64), 0, s, xs, inf, xext, count,
                       scale);
    return hipGetLastError();
}

hipError_t launch_table_build(uint32_t *table, uint32_t *zt, const uint32_t *xs, const uint32_t *inf,
                              unsigned base0, unsigned nb, unsigned win_bits, unsigned nwin, hipStream_t s) {
    const unsigned threads = nb * nwin;
    hipLaunchKernelGGL(k_table_build, dim3((threads + 63) / 64), dim3(64), 0, s, table, zt, xs, inf, base0, nb,
                       win_bits, nwin);
    return hipGetLastError();
}

hipError_t launch_toeplitz_fft_T(uint32_t *tcT, const uint32_t *poly, unsigned rows, unsigned N,
                                 const DevTables *T, hipStream_t s) {
    const unsigned blocks = rows * 16;
    switch (N) {
    case 1024: hipLaunchKernelGGL(k_toeplitz_fft_T<7>, dim3(blocks), dim3(64), 0, s, tcT, poly, rows, N, T); break;
    case 2048: hipLaunchKernelGGL(k_toeplitz_fft_T<8>, dim3(blocks), dim3(128), 0, s, tcT, poly, rows, N, T); break;
    case 4096: hipLaunchKernelGGL(k_toeplitz_fft_T<9>, dim3(blocks), dim3(256), 0, s, tcT, poly, rows, N, T); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_msm_comb(uint32_t *D, const uint32_t *tcT, const uint32_t *table, const uint32_t *inf,
                           unsigned rows, unsigned n2, unsigned win_bits, hipStream_t s) {
    const unsigned logn2 = n2 == 128 ? 7 : n2 == 256 ? 8 : 9;
    const dim3 grid(blocks_for(n2, rows));
    switch (win_bits) {
    case 8: hipLaunchKernelGGL(k_msm_comb<8>, grid, dim3(64), 0, s, D, tcT, table, inf, rows, n2, logn2); break;
    case 10: hipLaunchKernelGGL(k_msm_comb<10>, grid, dim3(64), 0, s, D, tcT, table, inf, rows, n2, logn2); break;
    case 12: hipLaunchKernelGGL(k_msm_comb<12>, grid, dim3(64), 0, s, D, tcT, table, inf, rows, n2, logn2); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_g1_transforms(uint32_t *D, unsigned rows, unsigned n2, const uint32_t *ops,
                                const uint32_t *index, hipStream_t s, LaunchHook hook) {
    const unsigned logn2 = n2 == 128 ? 7 : n2 == 256 ? 8 : 9;
    const unsigned nb = n2 / 2;
    for (unsigned st = 0; st + 1 < logn2; st++) {
        hipLaunchKernelGGL(k_dit_stage<TW_WNAF>, dim3(blocks_for(nb, rows)), dim3(64), 0, s, D, rows, st, ops, index);
        hook(LK_DIT);
    }
    hipLaunchKernelGGL(k_middle<TW_WNAF>, dim3(blocks_for(nb, rows)), dim3(64), 0, s, D, rows, logn2, ops, index);
    hook(LK_MIDDLE);
    for (unsigned st = 1; st < logn2; st++) {
        hipLaunchKernelGGL(k_dif_stage<TW_WNAF>, dim3(blocks_for(nb, rows)), dim3(64), 0, s, D, rows, logn2, st,
                           ops, index);
        hook(LK_DIF);
    }
    return hipGetLastError();
}

hipError_t launch_fft_output(uint32_t *h_fft, const uint32_t *D, unsigned rows, unsigned n2, hipStream_t s) {
    const unsigned logn2 = n2 == 128 ? 7 : n2 == 256 ? 8 : 9;
    hipLaunchKernelGGL(k_fft_output, dim3(blocks_for(n2, rows)), dim3(64), 0, s, h_fft, D, rows, logn2);
    return hipGetLastError();
}

}  // namespace fk20
