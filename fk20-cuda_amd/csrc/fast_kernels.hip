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
//  5. k_dit_stage / k_middle / k_dif_stage / k_twiddle_multi
//                      the G1 IFFT (DIT, bit-reversed in, natural out), the fused last IFFT
//                      stage + zero upper half + first FFT stage, and the G1 FFT (DIF,
//                      natural in, bit-reversed out), column-batched: a wave is one
//                      butterfly (or, in a twiddle layer, one point) for 64 rows, so the
//                      twiddle is wave-uniform and runs as a precomputed GLV + wNAF program
//                      (uniform control flow, no divergence).  Radix-2^2 schedule: the
//                      butterflies carry only w^+-128 twiddles, the point-wise layers the rest.
//  6. k_fft_output     D[pos][row] -> h_fft[row][bitrev(pos)] in the reference ABI.
// Same mathematics as fk20_msm.cu:23-66, g1p_fft.cu:25-330 and fk20_hext_fft2h_fft.cu:27-70;
// the layout and algorithms are this library's own.
#include <cstring>

#include "common.hip.h"
#include "fast.hip.h"
#include "kutil.hip.h"

namespace fk20 {

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

// ---------------------------------------------------------------- subgroup membership

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// A lane's point in LDS, packed (36 words, stride 37 to spread the lanes over the banks).
struct LdsSlot {
    uint32_t *w;
    __device__ void put(const g1p &p) {
        uint32_t t[36];
        g1_store_packed(t, p);
#pragma unroll
        for (int k = 0; k < 36; k++) w[k] = t[k];
    }
    __device__ void get(g1p &p) const {
        uint32_t t[36];
#pragma unroll
        for (int k = 0; k < 36; k++) t[k] = w[k];
        g1_load_packed(p, t);
    }
};

// flags[i / per_group] |= 1 when pts[i] lies outside the order-r subgroup.  The test's two
// saved points sit in LDS, which leaves room for two waves per SIMD without spills.
__global__ void __launch_bounds__(64, 2) k_subgroup_groups(uint32_t *flags, const uint32_t *pts, size_t count,
                                                           size_t per_group) {
    __shared__ uint32_t slots[2][64 * 37];
    const size_t i = (size_t)blockIdx.x * 64 + threadIdx.x;
    if (i >= count) return;
    uint32_t w[36];
    ld_words(w, pts + i * G1P_WORDS, 9);
    g1p p;
    g1_load_std(p, w);
    LdsSlot orig{slots[0] + threadIdx.x * 37}, base{slots[1] + threadIdx.x * 37};
    if (!g1_in_subgroup(p, orig, base)) atomicOr(flags + i / per_group, 1u);
}

// ---------------------------------------------------------------- rows outside G1
//
// On a point P outside the order-r subgroup, [a]P and [b]P differ whenever a != b as integers,
// even for a = b (mod r): a transform of such points is defined by its exact sequence of
// integer multipliers, not by the linear map mod r that the column-batched kernels compute.
// A row that holds such a point is recomputed here in the reference's order of operations:
// radix-2 decimation in time over bit-reversed positions, twiddle w^e applied to the right
// point of a butterfly as the canonical integer w^e mod r (g1p_fft.cu:25-170); the inverse
// with the butterflies transposed, w^-e applied after them and 2^-logn (and 2^-logn w^-128)
// in stage 1 (g1p_fft.cu:180-330); hext_fft -> h_fft as inverse, odd positions cleared,
// forward (fk20_hext_fft2h_fft.cu:27-70).  Products are double-and-add over complete
// formulas (g1_mul_std, as g1p_mul.cu:18).  One workgroup of 256 threads per row; the row's n
// points live in `work` (n * 36 words per row, row-contiguous); barriers order the stages.

__device__ __forceinline__ void ref_butterfly(uint32_t *Wr, unsigned s, unsigned t, bool inverse, unsigned logn,
                                              const DevTables *T) {
    const unsigned h = 1u << s, j = t & (h - 1);
    const unsigned l = ((t >> s) << (s + 1)) + j, r = l + h;
    const unsigned e = j << (8 - s);  // twiddle w^e of the order-512 root w
    g1p a, b;
    pt_load(a, Wr + (size_t)l * 36);
    pt_load(b, Wr + (size_t)r * 36);
    if (!inverse) {
        if (e) g1_mul_std(b, b, T->tw_std[e]);
        g1_addsub(a, b);
    } else {
        g1_addsub(a, b);
        if (s == 1) {
            g1_mul_std(a, a, T->inv2k_std[logn]);
            g1_mul_std(b, b, j ? T->inv2k_w384_std[logn] : T->inv2k_std[logn]);
        } else if (s > 1 && e) {
            g1_mul_std(b, b, T->tw_std[512 - e]);
        }
    }
    pt_store(Wr + (size_t)l * 36, a);
    pt_store(Wr + (size_t)r * 36, b);
}

// kind: 0 forward (in natural order -> out natural order), 1 inverse, 2 hext_fft -> h_fft.
// in_D != nullptr: the input is D[bitrev(j)][row] (packed Montgomery, the MSM's output layout)
// instead of in[row][j].
__global__ void __launch_bounds__(256) k_ref_rows(unsigned kind, uint32_t *out, const uint32_t *in,
                                                 const uint32_t *in_D, unsigned rows, unsigned logn, uint32_t *work,
                                                 const uint32_t *flags, unsigned flag_stride, const DevTables *T) {
    const unsigned row = blockIdx.x;
    if (!flags[(size_t)row * flag_stride]) return;  // uniform over the workgroup
    const unsigned n = 1u << logn, t = threadIdx.x;
    uint32_t *Wr = work + (size_t)row * n * 36;
    for (unsigned j = t; j < n; j += 256) {
        g1p p;
        if (in_D) {
            pt_load(p, in_D + ((size_t)bitrev(j, logn) * rows + row) * 36);
        } else {
            uint32_t w[36];
            ld_words(w, in + ((size_t)row * n + j) * G1P_WORDS, 9);
            g1_load_std(p, w);
        }
        pt_store(Wr + (size_t)(kind == 0 ? bitrev(j, logn) : j) * 36, p);
    }
    __syncthreads();
    if (kind != 0) {
#pragma unroll 1
        for (int s = (int)logn - 1; s >= 0; s--) {
            if (t < n / 2) ref_butterfly(Wr, (unsigned)s, t, true, logn, T);
            __syncthreads();
        }
        if (kind == 2) {
            // the upper half of the inverse's natural-order output sits at the odd positions
            for (unsigned j = 2 * t + 1; j < n; j += 512) {
                g1p z;
                g1_inf(z);
                pt_store(Wr + (size_t)j * 36, z);
            }
            __syncthreads();
        }
    }
    if (kind != 1) {
#pragma unroll 1
        for (unsigned s = 0; s < logn; s++) {
            if (t < n / 2) ref_butterfly(Wr, s, t, false, logn, T);
            __syncthreads();
        }
    }
    for (unsigned j = t; j < n; j += 256) {
        g1p p;
        pt_load(p, Wr + (size_t)(kind == 1 ? bitrev(j, logn) : j) * 36);
        uint32_t w[36];
        g1_store_std(w, p);
        st_words(out + ((size_t)row * n + j) * G1P_WORDS, w, 9);
    }
}

// ---------------------------------------------------------------- setup: window tables

// One thread per (base, window) of the launch's bases [base0, base0 + nb): entries d * B' for
// d = 1 .. H (B' = 2^(w win) B), built as projective multiples and converted to affine by
// Montgomery's batch trick in blocks of TABLE_BLK entries (one inversion per block).  zt holds
// (z_e, prefix_{e-1}) of the current block per thread, laid out [entry][thread] so that the
// 64 lanes of a wave touch consecutive words; it is sized by the launch, not by H, so every
// (base, window) of a setup fits in one or two launches.
__global__ void __launch_bounds__(64, 3) k_table_build(uint32_t *table, uint32_t *zt, const uint32_t *xs,
                                                    const uint32_t *inf, uint32_t *err, unsigned base0,
                                                    unsigned nb, unsigned win_bits, unsigned nwin) {
    const unsigned nthreads = nb * nwin;
    const unsigned t = blockIdx.x * 64 + threadIdx.x;
    if (t >= nthreads) return;
    const unsigned b = base0 + t / nwin, win = t % nwin;
    const unsigned H = 1u << (win_bits - 1), blk = min(H, TABLE_BLK);
    uint32_t *tab = table + ((size_t)b * nwin + win) * H * TABLE_ENTRY_WORDS;
    if (inf[b]) {
        for (unsigned e = 0; e < H; e++) {
            uint32_t zero[24] = {0};
            st_words(tab + (size_t)e * TABLE_ENTRY_WORDS, zero, 6);
        }
        return;
    }
    // B' = 2^(w win) B in affine form, so each entry is one mixed addition of B' (B' is not
    // infinity: the base is not, and E(Fp) has odd order).  An entry d B' is infinity only
    // for a base of small order outside G1; the block's inversion then fails and *err says so.
    g1p E;
    fe bx, by;
    {
        g1p B;
        pt_load(B, xs + (size_t)b * 36);
#pragma unroll 1
        for (unsigned d = 0; d < win_bits * win; d++) g1_dbl(B, B);
        g1_to_affine(bx, by, B);
    }
    E.x = bx;
    E.y = by;
    fe_one(E.z);
#pragma unroll 1
    for (unsigned e0 = 0; e0 < H; e0 += blk) {
        fe prefix;
        fe_one(prefix);
#pragma unroll 1
        for (unsigned i = 0; i < blk; i++) {
            const unsigned e = e0 + i;
            if (e) g1_add_aff(E, E, bx, by);
            uint32_t *z = zt + ((size_t)i * nthreads + t) * 24;
            fe_store(tab + (size_t)e * TABLE_ENTRY_WORDS, E.x);
            fe_store(tab + (size_t)e * TABLE_ENTRY_WORDS + 12, E.y);
            fe_store(z, E.z);
            fe_store(z + 12, prefix);  // prefix of z_e0 .. z_{e-1}
            fe_mul(prefix, prefix, E.z);
        }
        fe inv;
        fe_canonical(prefix, prefix);
        if (fe_is_zero_canonical(prefix)) atomicOr(err, 1u);
        fe_inv(inv, prefix);
#pragma unroll 1
        for (int i = (int)blk - 1; i >= 0; i--) {
            const unsigned e = e0 + (unsigned)i;
            const uint32_t *z = zt + ((size_t)i * nthreads + t) * 24;
            fe ze, pe, zinv, x, y;
            fe_load(ze, z);
            fe_load(pe, z + 12);
            fe_mul(zinv, inv, pe);
            fe_mul(inv, inv, ze);
            fe_load(x, tab + (size_t)e * TABLE_ENTRY_WORDS);
            fe_load(y, tab + (size_t)e * TABLE_ENTRY_WORDS + 12);
            fe_mul(x, x, zinv);
            fe_mul(y, y, zinv);
            fe_canonical(x, x);
            fe_canonical(y, y);
            fe_store(tab + (size_t)e * TABLE_ENTRY_WORDS, x);
            fe_store(tab + (size_t)e * TABLE_ENTRY_WORDS + 12, y);
        }
    }
}

// ---------------------------------------------------------------- Toeplitz + NTT (transposed)

__device__ __forceinline__ int toeplitz_src_T(unsigned i, unsigned j, unsigned N) {
    const unsigned k = N / 16;
    if (j == 0) return (int)(N - 1 - i);
    if (j <= k + 1) return -1;
    return (int)(31 - i + 16 * (j - (k + 2)));
}

// One (row, i) NTT of size n2 <= 512 per workgroup of 256 threads (n2/2 of them active).
__global__ void __launch_bounds__(256) k_toeplitz_fft_multi(Launch L, const DevTables *T) {
    __shared__ fs lds[512];
    unsigned local;
    const Item &it = L.item[find_item(L, blockIdx.x, local)];
    const unsigned logn = it.logn2, n = 1u << logn, half = n / 2, N = it.N;
    const unsigned t = threadIdx.x;
    const unsigned i = local % 16;
    const size_t row = local / 16;
    const uint32_t *poly = it.poly + row * N * FR_WORDS;
    if (t < half) {
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
            lds[bitrev(e, logn)] = a;
        }
    }
    __syncthreads();
    for (unsigned s = 0; s < logn; s++) {
        if (t < half) {
            const unsigned h = 1u << s, j = t & (h - 1);
            const unsigned l = ((t >> s) << (s + 1)) + j, r = l + h;
            fs x = lds[r], y = lds[l], w;
            fs_set(w, T->tw_mont[j << (8 - s)]);
            fs_mul(x, x, w);
            fs_add(lds[l], y, x);
            fs_sub(lds[r], y, x);
        }
        __syncthreads();
    }
    fs scale;
    fs_set(scale, T->inv2k_mont[logn]);
    if (t < half) {
        for (unsigned e = t; e < n; e += half) {
            fs a = lds[e];
            // the reference's order (k_ref_rows) applies 1/n2 inside the inverse transform
            if (it.scale_tc && !(it.rowflag && it.rowflag[(size_t)row * it.flag_stride])) fs_mul(a, a, scale);
            fs_from_mont(a, a);
            st_words(it.tcT + (((size_t)e * 16 + i) * it.rows + row) * FR_WORDS, a.l, 2);
        }
    }
}

// ---------------------------------------------------------------- window-table MSM

// acc += k * B for a standard-form scalar k < r (8 words, consumed) by signed fixed windows
// of WIN bits over B's table tb[win][d - 1] = d * 2^(WIN*win) * B (affine): one mixed
// addition per non-zero digit, no doublings.  k < r < 2^255 keeps the last window's carry
// inside the ceil(256/WIN) windows for every WIN in 8..13.
__device__ __forceinline__ void window_term(g1p &acc, uint32_t *k, const uint32_t *tb, unsigned WIN) {
    const unsigned NWIN = (256 + WIN - 1) / WIN, H = 1u << (WIN - 1), MASK = (1u << WIN) - 1;
    uint32_t carry = 0;
#pragma unroll 1
    for (unsigned win = 0; win < NWIN; win++) {
        const uint32_t bits = (k[0] & MASK) + carry;
        carry = bits > H ? 1u : 0u;
        const int d = (int)bits - (int)(carry << WIN);
        // shift the scalar right by WIN bits (WIN < 32)
#pragma unroll
        for (int q = 0; q < 7; q++) k[q] = (k[q] >> WIN) | (k[q + 1] << (32 - WIN));
        k[7] >>= WIN;
        if (d != 0) {
            const unsigned e = (unsigned)(d < 0 ? -d : d) - 1;
            const uint32_t *ent = tb + ((size_t)win * H + e) * TABLE_ENTRY_WORDS;
            fe x, y;
            fe_load(x, ent);
            fe_load(y, ent + 12);
            if (d < 0) fe_neg<2>(y, y);
            g1_add_aff(acc, acc, x, y);
        }
    }
}

// hext_fft[j] for 64 rows of one column j: sum over 16 terms of ceil(256/w) signed-window
// digits, one mixed addition each.  Output to D[bitrev(j)][row].
__global__ void __launch_bounds__(64) k_msm_multi(Launch L) {
    unsigned local;
    const Item &it = L.item[find_item(L, blockIdx.x, local)];
    const unsigned rows = it.rows, n2 = 1u << it.logn2, rg = (rows + 63) / 64;
    const unsigned WIN = it.win, NWIN = (256 + WIN - 1) / WIN, H = 1u << (WIN - 1);
    const unsigned j = local / rg;
    const unsigned row = min((local % rg) * 64 + threadIdx.x, rows - 1);  // spare lanes repeat the last row
    g1p acc;
    g1_inf(acc);
#pragma unroll 1
    for (unsigned i = 0; i < 16; i++) {
        const unsigned b = i * n2 + j;
        if (it.inf[b]) continue;
        uint32_t k[8];
        ld_words(k, it.tcT + (((size_t)j * 16 + i) * rows + row) * FR_WORDS, 2);
        window_term(acc, k, it.table + (size_t)b * NWIN * H * TABLE_ENTRY_WORDS, WIN);
    }
    pt_store(it.D + ((size_t)bitrev(j, it.logn2) * rows + row) * 36, acc);
}

// The same MSM for launches whose items all have at most MSM_SMALL_ROWS rows, where the
// 16 x ceil(256/w) mixed additions of a lane are latency: lanes are (term i, row) pairs, 16
// terms x 4 rows per wave, each lane sums one term's windows, and the 16 partial sums of a
// row are added in LDS (4 levels).  Output to D[bitrev(j)][row] like k_msm_multi.
static constexpr unsigned MSM_SMALL_ROWS = 16;

__global__ void __launch_bounds__(64) k_msm_multi_small(Launch L) {
    __shared__ __attribute__((aligned(16))) uint32_t lds[64 * 36];
    unsigned local;
    const Item &it = L.item[find_item(L, blockIdx.x, local)];
    const unsigned rows = it.rows, n2 = 1u << it.logn2, rg4 = (rows + 3) / 4;
    const unsigned WIN = it.win, NWIN = (256 + WIN - 1) / WIN, H = 1u << (WIN - 1);
    const unsigned j = local / rg4, t = threadIdx.x, i = t % 16, r = t / 16;
    const unsigned row0 = (local % rg4) * 4 + r, row = min(row0, rows - 1);  // spare lanes repeat a row
    const unsigned b = i * n2 + j;
    g1p acc;
    g1_inf(acc);
    if (!it.inf[b]) {
        uint32_t k[8];
        ld_words(k, it.tcT + (((size_t)j * 16 + i) * rows + row) * FR_WORDS, 2);
        window_term(acc, k, it.table + (size_t)b * NWIN * H * TABLE_ENTRY_WORDS, WIN);
    }
    pt_store(lds + t * 36, acc);
#pragma unroll 1
    for (unsigned h = 8; h >= 1; h /= 2) {
        __syncthreads();
        g1p o;
        if (i < h) {
            pt_load(o, lds + (t + h) * 36);
            g1_add(acc, acc, o);
        }
        __syncthreads();
        if (i < h) pt_store(lds + t * 36, acc);
    }
    if (i == 0 && row0 < rows) pt_store(it.D + ((size_t)bitrev(j, it.logn2) * rows + row) * 36, acc);
}

// ---------------------------------------------------------------- KZG commitments
//
// commit[row] = sum_i poly[row][i] * srs[i] (FK20Py kzg_proofs.py:159 commit_to_poly) over
// window tables of the SRS points built like the FK20 setup's.  Lanes are rows; a block sums
// one chunk of consecutive coefficients into partial[chunk][row], then two reduction passes
// add the chunks.

__global__ void __launch_bounds__(64) k_commit_partial(const uint32_t *poly, const uint32_t *table,
                                                       const uint32_t *inf, uint32_t *partial, unsigned rows,
                                                       unsigned N, unsigned win_bits, unsigned chunk) {
    const unsigned rg = (rows + 63) / 64, c = blockIdx.x / rg;
    const unsigned row = min((blockIdx.x % rg) * 64 + threadIdx.x, rows - 1);  // spare lanes repeat the last row
    const unsigned NWIN = (256 + win_bits - 1) / win_bits, H = 1u << (win_bits - 1);
    const unsigned i0 = c * chunk, i1 = min(i0 + chunk, N);
    g1p acc;
    g1_inf(acc);
#pragma unroll 1
    for (unsigned i = i0; i < i1; i++) {
        if (inf[i]) continue;
        // the caller's coefficient may be any 256-bit representative: reduce it below r, so
        // the top signed window's carry stays inside the ceil(256/w) windows (at w = 8 the
        // windows end exactly at bit 256, and a value >= 2^255 would carry out of them)
        fs k;
        ld_words(k.l, poly + ((size_t)row * N + i) * FR_WORDS, 2);
        fs_reduce_std(k);
        window_term(acc, k.l, table + (size_t)i * NWIN * H * TABLE_ENTRY_WORDS, win_bits);
    }
    pt_store(partial + ((size_t)c * rows + row) * 36, acc);
}

// out[q][row] = sum of in[q*group .. min((q+1)*group, nparts)][row]; with std_out the single
// group's sum goes to out[row] as a canonical standard-form g1p_t.
__global__ void __launch_bounds__(64) k_commit_reduce(const uint32_t *in, uint32_t *out, unsigned rows,
                                                      unsigned nparts, unsigned group, unsigned nout,
                                                      unsigned std_out) {
    const size_t t = (size_t)blockIdx.x * 64 + threadIdx.x;
    if (t >= (size_t)rows * nout) return;
    const unsigned row = (unsigned)(t % rows), q = (unsigned)(t / rows);
    const unsigned p0 = q * group, p1 = min(p0 + group, nparts);
    g1p acc, x;
    g1_inf(acc);
#pragma unroll 1
    for (unsigned p = p0; p < p1; p++) {
        pt_load(x, in + ((size_t)p * rows + row) * 36);
        g1_add(acc, acc, x);
    }
    if (std_out) {
        uint32_t w[36];
        g1_store_std(w, acc);
        st_words(out + (size_t)row * G1P_WORDS, w, 9);
    } else {
        pt_store(out + ((size_t)q * rows + row) * 36, acc);
    }
}

// ---------------------------------------------------------------- G1 transform stages
//
// The butterfly, twiddle-layer, persistent-chain and drop-in load kernels are templates over
// a schedule variant (stage_kernels.hip.h), built once per variant (stage_variant.hip);
// launches go through the variant's StageKernels table.

extern const StageKernels stage_kernels_3_0_2, stage_kernels_3_0_3, stage_kernels_3_1_2, stage_kernels_3_1_3,
    stage_kernels_4_0_2, stage_kernels_4_0_3, stage_kernels_4_1_2, stage_kernels_4_1_3, stage_kernels_5_0_2,
    stage_kernels_5_0_3, stage_kernels_5_1_2, stage_kernels_5_1_3;

const StageKernels *stage_kernels(unsigned wnaf, bool r22, unsigned occupancy) {
    static const StageKernels *const table[3][2][2] = {
        {{&stage_kernels_3_0_2, &stage_kernels_3_0_3}, {&stage_kernels_3_1_2, &stage_kernels_3_1_3}},
        {{&stage_kernels_4_0_2, &stage_kernels_4_0_3}, {&stage_kernels_4_1_2, &stage_kernels_4_1_3}},
        {{&stage_kernels_5_0_2, &stage_kernels_5_0_3}, {&stage_kernels_5_1_2, &stage_kernels_5_1_3}}};
    if (wnaf < 3 || wnaf > 5 || occupancy < 2 || occupancy > 3) return nullptr;
    return table[wnaf - 3][r22 ? 1 : 0][occupancy - 2];
}
//
// Clears the chain's flags of every item and its work counter / error word (ctl[0..1]).
__global__ void k_chain_reset(Chain C) {
    if (blockIdx.x == 0 && threadIdx.x < 2) C.ctl[threadIdx.x] = 0;
    const Item &it = C.L.item[blockIdx.x];
    const unsigned n = ((it.rows + 63) / 64) << it.logn2;
    for (unsigned e = threadIdx.x; e < n; e += blockDim.x) it.flags[e] = 0;
}

// D[pos][row] -> the item's output format.  D is in bit-reversed order, so the natural
// index of pos is bitrev(pos) and the data-availability order (FK20Py
// list_to_reverse_bit_order of the natural list) is pos itself.
//   FK20_OUT_G1P / _DA:               g1p_t, canonical standard projective (144 bytes)
//   FK20_OUT_COMPRESSED / _DA:        48-byte compressed G1 (g1_compress)
__global__ void __launch_bounds__(64) k_output_multi(Launch L) {
    unsigned local;
    const Item &it = L.item[find_item(L, blockIdx.x, local)];
    const unsigned rows = it.rows, rg = (rows + 63) / 64, logn2 = it.logn2;
    const unsigned pos = local / rg;
    const unsigned row = (local % rg) * 64 + threadIdx.x;
    if (row >= rows) return;
    if (it.rowflag && it.rowflag[(size_t)row * it.flag_stride]) return;  // written by k_ref_rows
    g1p p;
    pt_load(p, it.D + ((size_t)pos * rows + row) * 36);
    const unsigned fmt = it.out_format;
    const size_t slot = (size_t)row * (1u << logn2) + ((fmt & 1u) ? pos : bitrev(pos, logn2));
    if (fmt & 2u) {
        uint32_t w[12];
        g1_compress(w, p);
        st_words(it.h_fft + slot * 12, w, 3);
    } else {
        uint32_t w[36];
        g1_store_std(w, p);
        st_words(it.h_fft + slot * G1P_WORDS, w, 9);
    }
}

// ---------------------------------------------------------------- launchers

hipError_t launch_subgroup_groups(uint32_t *flags, const uint32_t *pts, size_t count, size_t per_group,
                                  hipStream_t s) {
    if (!count || !per_group) return hipSuccess;
    const size_t groups = (count + per_group - 1) / per_group;
    // a fill kernel, not hipMemsetAsync: callers may capture these launches into a graph
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((groups + 255) / 256)), dim3(256), 0, s, flags, groups, 0u);
    hipLaunchKernelGGL(k_subgroup_groups, dim3((unsigned)((count + 63) / 64)), dim3(64), 0, s, flags, pts, count,
                       per_group);
    return hipGetLastError();
}

hipError_t launch_scale_bases(uint32_t *xs, uint32_t *inf, const uint32_t *xext, size_t count,
                              const uint32_t *scale, hipStream_t s) {
    hipLaunchKernelGGL(k_scale_bases, dim3((unsigned)((count + 63) / 64)), dim3(64), 0, s, xs, inf, xext, count,
                       scale);
    return hipGetLastError();
}

hipError_t launch_table_build(uint32_t *table, uint32_t *zt, const uint32_t *xs, const uint32_t *inf,
                              uint32_t *err, unsigned base0, unsigned nb, unsigned win_bits, unsigned nwin,
                              hipStream_t s) {
    const unsigned threads = nb * nwin;
    hipLaunchKernelGGL(k_table_build, dim3((threads + 63) / 64), dim3(64), 0, s, table, zt, xs, inf, err, base0,
                       nb, win_bits, nwin);
    return hipGetLastError();
}

// Commitment chunking: coefficients per partial sum, partials per first reduction group.
static constexpr unsigned COMMIT_CHUNK = 16, COMMIT_GROUP = 16;

size_t commit_workspace_bytes(unsigned rows, unsigned N) {
    const size_t nparts = (N + COMMIT_CHUNK - 1) / COMMIT_CHUNK;
    const size_t nmid = (nparts + COMMIT_GROUP - 1) / COMMIT_GROUP;
    return (nparts + nmid) * rows * 144;
}

hipError_t launch_commit(uint32_t *out, const uint32_t *poly, const uint32_t *table, const uint32_t *inf,
                         uint32_t *work, unsigned rows, unsigned N, unsigned win_bits, hipStream_t s) {
    if (!rows) return hipSuccess;
    const unsigned nparts = (N + COMMIT_CHUNK - 1) / COMMIT_CHUNK;
    const unsigned nmid = (nparts + COMMIT_GROUP - 1) / COMMIT_GROUP;
    const unsigned rg = (rows + 63) / 64;
    uint32_t *partial = work, *mid = work + (size_t)nparts * rows * 36;
    hipLaunchKernelGGL(k_commit_partial, dim3(nparts * rg), dim3(64), 0, s, poly, table, inf, partial, rows, N,
                       win_bits, COMMIT_CHUNK);
    const size_t t1 = (size_t)rows * nmid;
    hipLaunchKernelGGL(k_commit_reduce, dim3((unsigned)((t1 + 63) / 64)), dim3(64), 0, s, partial, mid, rows, nparts,
                       COMMIT_GROUP, nmid, 0u);
    hipLaunchKernelGGL(k_commit_reduce, dim3((rows + 63) / 64), dim3(64), 0, s, mid, out, rows, nmid, nmid, 1u, 1u);
    return hipGetLastError();
}

// Fills L.blk_end from per-item block counts; returns the total.
static unsigned set_blocks(Launch &L, const unsigned *per_item) {
    unsigned total = 0;
    for (unsigned k = 0; k < L.nitems; k++) {
        total += per_item[k];
        L.blk_end[k] = total;
    }
    return total;
}

// Resident waves of k_chain: two per SIMD on 256 CUs.
static constexpr unsigned CHAIN_WAVES = 2048;

hipError_t launch_batch(const Launch &base, const DevTables *T, const uint32_t *ops, const uint32_t *index,
                        const StageKernels &K, hipStream_t s, LaunchHook hook, uint32_t *chain,
                        uint64_t chain_wait) {
    if (chain && !K.chain) return hipErrorInvalidValue;
    Launch L = base;
    unsigned per[MAX_ITEMS], lmax = 0;
    for (unsigned k = 0; k < L.nitems; k++) lmax = std::max(lmax, L.item[k].logn2);
    if (chain) {
        Chain C;
        memset(&C, 0, sizeof C);
        C.L = L;
        C.ctl = chain;
        hipLaunchKernelGGL(k_chain_reset, dim3(L.nitems), dim3(256), 0, s, C);
    }
    // Toeplitz + NTT
    for (unsigned k = 0; k < L.nitems; k++) per[k] = L.item[k].rows * 16;
    unsigned total = set_blocks(L, per);
    if (total) hipLaunchKernelGGL(k_toeplitz_fft_multi, dim3(total), dim3(256), 0, s, L, T);
    hook(LK_TOEPLITZ_FFT);
    // MSM
    bool small = true;
    for (unsigned k = 0; k < L.nitems; k++) small = small && L.item[k].rows <= MSM_SMALL_ROWS;
    for (unsigned k = 0; k < L.nitems; k++)
        per[k] = (1u << L.item[k].logn2) * (small ? (L.item[k].rows + 3) / 4 : (L.item[k].rows + 63) / 64);
    total = set_blocks(L, per);
    if (total && small) hipLaunchKernelGGL(k_msm_multi_small, dim3(total), dim3(64), 0, s, L);
    else if (total) hipLaunchKernelGGL(k_msm_multi, dim3(total), dim3(64), 0, s, L);
    hook(LK_MSM);
    // transform stages: global stage t = 0 .. 2*lmax - 2; item k (L = logn2) runs local stage
    // u = t - (lmax - L) while 0 <= u <= 2L - 2 (DIT right-aligned to the middle at t = lmax-1,
    // DIF left-aligned after it)
    const unsigned nt = 2 * lmax - 1;
    if (chain) {
        Chain C;
        memset(&C, 0, sizeof C);
        C.L = L;
        C.ctl = chain;
        C.wait_ticks = chain_wait;
        C.nstages = nt;
        unsigned acc = 0;
        for (unsigned t = 0; t < nt; t++) {
            unsigned in_stage = 0;
            for (unsigned k = 0; k < L.nitems; k++) {
                const Item &it = L.item[k];
                const int u = t < lmax - 1 ? (int)t - (int)(lmax - it.logn2)
                            : (int)it.logn2 - 1 + (int)(t - (lmax - 1));
                const bool live = t < lmax - 1 ? u >= 0 : u <= (int)(2 * it.logn2 - 2);
                C.ustage[t][k] = (int8_t)(live ? u : -1);
                in_stage += live ? (1u << (it.logn2 - 1)) * ((it.rows + 63) / 64) : 0;
                C.blk_end[t][k] = in_stage;
            }
            acc += in_stage;
            C.t_end[t] = acc;
        }
        C.total = acc;
        if (acc) hipLaunchKernelGGL(K.chain, dim3(std::min(acc, CHAIN_WAVES)), dim3(64), 0, s, C, ops, index);
        hook(LK_CHAIN);
    } else {
        // one launch per stage; IFFT DIT stages right-aligned to the middle launch.  In the
        // radix-2^2 schedule a DIT stage whose transposed DIF stage is odd is preceded by that
        // stage's twiddle layer.
        for (unsigned t = 0; t + 1 < lmax; t++) {
            for (unsigned k = 0; k < L.nitems; k++) {
                const Item &it = L.item[k];
                const int st = (int)t - (int)(lmax - it.logn2);
                const unsigned sd = it.logn2 - 1 - (unsigned)st;
                const bool layer = K.r22 && st >= 0 && (sd & 1) && layer_count(sd, it.logn2);
                L.item[k].stage = layer ? (int)sd : -1;
                per[k] = layer ? layer_count(sd, it.logn2) * ((it.rows + 63) / 64) : 0;
            }
            total = set_blocks(L, per);
            if (total) {
                hipLaunchKernelGGL(K.twiddle, dim3(total), dim3(64), 0, s, L, ops, index, (unsigned)TW_INV);
                hook(LK_TWIDDLE);
            }
            for (unsigned k = 0; k < L.nitems; k++) {
                const int st = (int)t - (int)(lmax - L.item[k].logn2);
                L.item[k].stage = st;
                per[k] = st >= 0 ? (1u << (L.item[k].logn2 - 1)) * ((L.item[k].rows + 63) / 64) : 0;
            }
            total = set_blocks(L, per);
            if (total) hipLaunchKernelGGL(K.dit, dim3(total), dim3(64), 0, s, L, ops, index);
            hook(LK_DIT);
        }
        for (unsigned k = 0; k < L.nitems; k++) {
            L.item[k].stage = 0;
            per[k] = (1u << (L.item[k].logn2 - 1)) * ((L.item[k].rows + 63) / 64);
        }
        total = set_blocks(L, per);
        if (total) hipLaunchKernelGGL(K.middle, dim3(total), dim3(64), 0, s, L, ops, index);
        hook(LK_MIDDLE);
        // FFT DIF stages 1.., left-aligned after the middle launch
        for (unsigned t = 1; t < lmax; t++) {
            for (unsigned k = 0; k < L.nitems; k++) {
                const int st = t < L.item[k].logn2 ? (int)t : -1;
                L.item[k].stage = st;
                per[k] = st >= 0 ? (1u << (L.item[k].logn2 - 1)) * ((L.item[k].rows + 63) / 64) : 0;
            }
            total = set_blocks(L, per);
            if (total) hipLaunchKernelGGL(K.dif, dim3(total), dim3(64), 0, s, L, ops, index);
            hook(LK_DIF);
            // even DIF stages (radix-2^2): their w^128 twiddles, on the second half's outputs
            for (unsigned k = 0; k < L.nitems; k++) {
                const Item &it = L.item[k];
                const bool q = K.r22 && t + 1 < it.logn2 && (t & 1) == 0;
                L.item[k].stage = q ? (int)t : -1;
                per[k] = q ? (1u << (it.logn2 - 2)) * ((it.rows + 63) / 64) : 0;
            }
            total = set_blocks(L, per);
            if (total) {
                hipLaunchKernelGGL(K.quarter, dim3(total), dim3(64), 0, s, L);
                hook(LK_DIF);
            }
            // odd DIF stages: their twiddle layer (radix-2^2)
            for (unsigned k = 0; k < L.nitems; k++) {
                const Item &it = L.item[k];
                const bool layer = K.r22 && t < it.logn2 && (t & 1) && layer_count(t, it.logn2);
                L.item[k].stage = layer ? (int)t : -1;
                per[k] = layer ? layer_count(t, it.logn2) * ((it.rows + 63) / 64) : 0;
            }
            total = set_blocks(L, per);
            if (total) {
                hipLaunchKernelGGL(K.twiddle, dim3(total), dim3(64), 0, s, L, ops, index, (unsigned)TW_FWD);
                hook(LK_TWIDDLE);
            }
        }
    }
    // output
    for (unsigned k = 0; k < L.nitems; k++) per[k] = (1u << L.item[k].logn2) * ((L.item[k].rows + 63) / 64);
    total = set_blocks(L, per);
    if (total) hipLaunchKernelGGL(k_output_multi, dim3(total), dim3(64), 0, s, L);
    hook(LK_OUTPUT);
    return hipGetLastError();
}

// The drop-in transforms as one item of the batched launches: the DIT stage kernel runs
// IFFT stages 0 .. L-1 (the last one through the middle's positions), the DIF kernel FFT
// stages 0 .. L-1 (the first likewise), and the output kernel writes natural order from the
// DIF's bit-reversed positions (format 0) or position by position (format 1).
hipError_t launch_plain(PlainKind kind, uint32_t *out, const uint32_t *in, const uint32_t *xpk, unsigned rows,
                        unsigned logn, uint32_t *D, uint32_t *tcT, int8_t *digits, uint32_t *flags,
                        const DevTables *T, const Programs &P, hipStream_t s) {
    if (!rows) return hipSuccess;
    // rows holding a point outside G1 are recomputed by k_ref_rows: the transform kinds check
    // their inputs here, one flag per row; PLAIN_POLY2H_FFT's caller has set flags[0] from the
    // bases, which every row's MSM output combines
    if (kind != PLAIN_POLY2H_FFT) {
        hipError_t e = launch_subgroup_groups(flags, in, (size_t)rows << logn, (size_t)1 << logn, s);
        if (e != hipSuccess) return e;
    }
    Launch L;
    memset(&L, 0, sizeof L);
    L.nitems = 1;
    Item &it = L.item[0];
    it.poly = in;
    it.h_fft = out;
    it.tcT = tcT;
    it.D = D;
    it.rows = rows;
    it.N = 8u << logn;
    it.logn2 = logn;
    it.scale_tc = 1;
    it.rowflag = flags;
    it.flag_stride = kind == PLAIN_POLY2H_FFT ? 0 : 1;
    const unsigned rg = (rows + 63) / 64, n = 1u << logn;
    unsigned per[MAX_ITEMS] = {};
    const uint32_t *scale = P.ops + P.scale_off[logn];
    const StageKernels &K = default_stage_kernels();  // radix-2^2, built for P's width
    // radix-2^2 twiddle layers (see stage_positions): DIF stage t's after its butterflies,
    // the transposed DIF stage's before DIT stage t
    auto layer = [&](unsigned sd, unsigned dir) {
        const unsigned cnt = (sd & 1) ? layer_count(sd, logn) : 0;
        if (!cnt) return;
        it.stage = (int)sd;
        per[0] = cnt * rg;
        hipLaunchKernelGGL(K.twiddle, dim3(set_blocks(L, per)), dim3(64), 0, s, L, P.ops, P.index, dir);
    };
    auto stage = [&](void (*kernel)(Launch, const uint32_t *, const uint32_t *), int st) {
        const bool dit = kernel == K.dit, dif = kernel == K.dif;
        if (dit) layer(logn - 1 - (unsigned)st, (unsigned)TW_INV);
        it.stage = st;
        per[0] = n / 2 * rg;
        const unsigned total = set_blocks(L, per);
        hipLaunchKernelGGL(kernel, dim3(total), dim3(64), 0, s, L, P.ops, P.index);
        if (dif && (st & 1) == 0 && (unsigned)st + 1 < logn) {  // its w^128 twiddles
            per[0] = (n / 4) * rg;
            hipLaunchKernelGGL(K.quarter, dim3(set_blocks(L, per)), dim3(64), 0, s, L);
        }
        if (dif) layer((unsigned)st, (unsigned)TW_FWD);
    };
    switch (kind) {
    case PLAIN_G1_FFT:
        hipLaunchKernelGGL(K.plain_load, dim3(n * rg), dim3(64), 0, s, D, in, rows, logn, 0u,
                           (const uint32_t *)nullptr);
        for (unsigned t = 0; t < logn; t++) stage(K.dif, (int)t);
        break;
    case PLAIN_G1_IFT:
        hipLaunchKernelGGL(K.plain_load, dim3(n * rg), dim3(64), 0, s, D, in, rows, logn, 1u, scale);
        for (unsigned t = 0; t < logn; t++) stage(K.dit, (int)t);
        it.out_format = 1;
        break;
    case PLAIN_HEXT_FFT2H_FFT:
    case PLAIN_POLY2H_FFT:
        if (kind == PLAIN_HEXT_FFT2H_FFT) {
            hipLaunchKernelGGL(K.plain_load, dim3(n * rg), dim3(64), 0, s, D, in, rows, logn, 1u, scale);
        } else {
            per[0] = rows * 16;
            hipLaunchKernelGGL(k_toeplitz_fft_multi, dim3(set_blocks(L, per)), dim3(256), 0, s, L, T);
            hipError_t e = launch_msm_plain_T(D, tcT, xpk, digits, rows, n, s);
            if (e != hipSuccess) return e;
            // rows of a setup outside G1 take the reference's order from the MSM output on,
            // with tcT (no longer read) as their workspace
            hipLaunchKernelGGL(k_ref_rows, dim3(rows), dim3(256), 0, s, 2u, out, (const uint32_t *)nullptr,
                               (const uint32_t *)D, rows, logn, tcT, (const uint32_t *)flags, 0u, T);
        }
        for (unsigned t = 0; t + 1 < logn; t++) stage(K.dit, (int)t);
        stage(K.middle, 0);
        for (unsigned t = 1; t < logn; t++) stage(K.dif, (int)t);
        break;
    }
    per[0] = n * rg;
    hipLaunchKernelGGL(k_output_multi, dim3(set_blocks(L, per)), dim3(64), 0, s, L);
    // flagged rows of the transform kinds: their input rows are intact (the output kernel
    // skipped them, so an aliased output has not overwritten them) and D is free as workspace
    if (kind != PLAIN_POLY2H_FFT)
        hipLaunchKernelGGL(k_ref_rows, dim3(rows), dim3(256), 0, s,
                           kind == PLAIN_G1_FFT ? 0u : kind == PLAIN_G1_IFT ? 1u : 2u, out, in,
                           (const uint32_t *)nullptr, rows, logn, D, (const uint32_t *)flags, 1u, T);
    return hipGetLastError();
}

}  // namespace fk20
