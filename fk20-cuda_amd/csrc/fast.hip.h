// Batched FK20 path: twiddle-program encoding and launchers (fast_kernels.hip, ctx.cpp).
#pragma once
#include "common.hip.h"
#include "ladder.hip.h"
#include "schedule.hip.h"

namespace fk20 {

// Twiddle programs: op encoding and the ladder that runs them in ladder.hip.h.
// direction rows of the program index: index[dir * 512 + idx] = offset of w^(+-idx)'s program
static constexpr int TW_FWD = 0;
static constexpr int TW_INV = 1;
// index[TW_SCALE + j] = offset of the program of 2^-j (j < 10), the 1/n of the inverse transforms
static constexpr int TW_SCALE = 2 * 512;
// index[TW_LAYER + layer_table_offset(logn2, sd, dir) + k], k < layer_count(sd, logn2): the
// k-th item of the twiddle layer after odd DIF stage sd, packed q | e << 16 (the LayerTw of
// layer_twiddle), in order of decreasing ladder cost of its program in direction dir.  A
// layer launch hands its items to the SIMDs in that order: the long ladders spread over the
// chip first and the short ones (w^+-128, w^256) fill in after them, instead of two long
// ladders meeting on one SIMD while others run short ones.
static constexpr int TW_LAYER = TW_SCALE + 10;
static constexpr int TW_LAYER_SIZE = 4096;  // 2 x (225 + 513 + 1281) items for n2 = 128, 256, 512
static constexpr int TW_INDEX_SIZE = TW_LAYER + TW_LAYER_SIZE;

FK20_HD unsigned layer_table_offset(unsigned logn2, unsigned sd, unsigned dir) {
    unsigned off = 0;
    for (unsigned L = 7; L <= 9; L++)
        for (unsigned s = 1; s < L; s += 2)
            for (unsigned d = 0; d < 2; d++) {
                if (L == logn2 && s == sd && d == dir) return off;
                off += layer_count(s, L);
            }
    return off;
}

// Twiddle programs of one device (ctx.cpp): ops and index in device memory; the scale
// programs' offsets also on the host, so launchers can pass them as plain pointers.
struct Programs {
    uint32_t *ops = nullptr;
    uint32_t *index = nullptr;
    uint32_t scale_off[10] = {};
    bool owned = false;  // ops and index were allocated by build_programs
    unsigned wnaf = 0;   // digit width the programs were built for
};
// Builds and checks every program on the host (each evaluates to its scalar over Fr with
// phi = lambda), then copies them to new allocations on the current device.
hipError_t build_programs(Programs &P, unsigned wnaf = TW_WNAF);
// The same, copied into caller-owned device memory (ops_cap words of ops, TW_INDEX_SIZE
// words of index); no allocation.
hipError_t load_programs(Programs &P, uint32_t *ops_dev, size_t ops_cap, uint32_t *index_dev);
void free_programs(Programs &P);

hipError_t launch_scale_bases(uint32_t *xs, uint32_t *inf, const uint32_t *xext, size_t count,
                              const uint32_t *scale, hipStream_t s);
// flags[g] = 1 if any of the points pts[g * per_group .. (g + 1) * per_group) (standard-form
// g1p_t, count in all) lies outside the order-r subgroup, else 0 (the flags of the groups are
// cleared first); groups = ceil(count / per_group).
hipError_t launch_subgroup_groups(uint32_t *flags, const uint32_t *pts, size_t count, size_t per_group,
                                  hipStream_t s);
// Words per window-table entry: the affine (x, y), 24 words, padded to 32 so that an entry is
// one 128-byte line of the caches (a 96-byte entry straddles two lines in half the cases).
static constexpr unsigned TABLE_ENTRY_WORDS = 32;
// Window tables of bases [base0, base0 + nb): zt is workspace of table_build_scratch_bytes(nb,
// win_bits, nwin) bytes.
static constexpr unsigned TABLE_BLK = 256;  // entries per batch inversion
inline size_t table_build_scratch_bytes(unsigned nb, unsigned win_bits, unsigned nwin) {
    const unsigned H = 1u << (win_bits - 1);
    return (size_t)nb * nwin * (H < TABLE_BLK ? H : TABLE_BLK) * 96;
}
// *err |= 1 when an entry d * 2^(w win) * base is infinity (a base of small order outside G1:
// an affine table cannot hold it).
hipError_t launch_table_build(uint32_t *table, uint32_t *zt, const uint32_t *xs, const uint32_t *inf,
                              uint32_t *err, unsigned base0, unsigned nb, unsigned win_bits, unsigned nwin,
                              hipStream_t s);
// KZG commitments of `rows` polynomials of length N over an SRS setup's window tables.
size_t commit_workspace_bytes(unsigned rows, unsigned N);
hipError_t launch_commit(uint32_t *out, const uint32_t *poly, const uint32_t *table, const uint32_t *inf,
                         uint32_t *work, unsigned rows, unsigned N, unsigned win_bits, hipStream_t s);
// One group of rows sharing a setup (hence N) and buffers; the unit of a batched launch.
static constexpr unsigned MAX_ITEMS = 16;
struct Item {
    const uint32_t *poly;   // [rows][N] fr_t
    uint32_t *h_fft;        // [rows][n2] g1p_t
    uint32_t *tcT;          // workspace [n2][16][rows] fr_t
    uint32_t *D;            // workspace [n2][rows] packed Montgomery points
    uint32_t *flags;        // workspace [ceil(rows/64)][n2]: last local stage written (chain)
    const uint32_t *table;  // setup window tables
    const uint32_t *inf;    // setup infinity flags
    uint32_t rows, N, logn2, win;
    uint32_t out_format;    // FK20_OUT_* (bit 0: data-availability order, bit 1: compressed)
    uint32_t scale_tc;      // 1: the NTT output is scaled by 1/n2 (setups without folded 1/n2)
    int32_t stage;          // set per launch by launch_batch
    // rowflag[row * flag_stride] != 0: the row holds a point outside the order-r subgroup and
    // is recomputed by launch_ref_rows, so the output kernel leaves it alone (nullptr: none)
    const uint32_t *rowflag;
    uint32_t flag_stride;
};
struct Launch {
    Item item[MAX_ITEMS];
    uint32_t blk_end[MAX_ITEMS];  // per-item block prefix sums of the current launch
    uint32_t nitems;
};

// The transform stages of a batch as one persistent launch (k_chain): work items are
// numbered stage by stage; t_end[t] = items in global stages 0..t, blk_end[t][k] = items of
// items 0..k within stage t, ustage[t][k] = item k's local stage at global stage t.
static constexpr unsigned MAX_STAGES = 17;  // 2 * 9 - 1 for n2 = 512
struct Chain {
    Launch L;
    uint32_t *ctl;  // [0] work counter, [1] error word (a bounded wait expired)
    uint64_t wait_ticks;  // bound of each wait, in ticks of the constant-rate wall clock
    uint32_t total, nstages;
    uint32_t t_end[MAX_STAGES];
    uint32_t blk_end[MAX_STAGES][MAX_ITEMS];
    int8_t ustage[MAX_STAGES][MAX_ITEMS];
};

// Kernel kinds of the batched path, reported by the per-launch timing of the context.
enum LaunchKind {
    LK_TOEPLITZ_FFT = 0, LK_MSM = 1, LK_DIT = 2, LK_MIDDLE = 3, LK_DIF = 4, LK_OUTPUT = 5, LK_CHAIN = 6,
    LK_TWIDDLE = 7
};
// Called after each launch when the context times launches (fn == nullptr otherwise).
struct LaunchHook {
    void (*fn)(void *arg, int kind);
    void *arg;
    void operator()(int kind) const {
        if (fn) fn(arg, kind);
    }
};

// The G1 transform kernels of one schedule variant (stage_kernels.hip.h): wNAF width of the
// twiddle programs (3..5), radix-2^2 (r22) or radix-2 butterflies, waves per SIMD the
// register allocation leaves room for (2 or 3).
// chain is set for the radix-2 variants only (the persistent form runs radix-2 stages).
struct StageKernels {
    void (*dit)(Launch, const uint32_t *, const uint32_t *) = nullptr;
    void (*middle)(Launch, const uint32_t *, const uint32_t *) = nullptr;
    void (*dif)(Launch, const uint32_t *, const uint32_t *) = nullptr;
    void (*twiddle)(Launch, const uint32_t *, const uint32_t *, unsigned) = nullptr;
    // radix-2^2 only: the w^128 twiddles of a DIF stage, after its butterflies (k_quarter_multi)
    void (*quarter)(Launch) = nullptr;
    void (*chain)(Chain, const uint32_t *, const uint32_t *) = nullptr;
    void (*plain_load)(uint32_t *, const uint32_t *, unsigned, unsigned, unsigned, const uint32_t *) = nullptr;
    int wnaf = 0;
    bool r22 = false;
    int occupancy = 0;
};
// The built variant, or nullptr: wnaf 3..5, r22, waves per SIMD 2..3.
const StageKernels *stage_kernels(unsigned wnaf, bool r22, unsigned occupancy);
// The default schedule: the drop-in entry points and new contexts use it.
static constexpr unsigned DEFAULT_OCCUPANCY = 2;
inline const StageKernels &default_stage_kernels() { return *stage_kernels(TW_WNAF, true, DEFAULT_OCCUPANCY); }

// The whole batched path for the items of `L`: Toeplitz + NTT, MSM, 2*Lmax - 1 transform
// stages, output -- one launch per stage covering every item.
// chain != nullptr: the transform stages run as one persistent k_chain launch whose work
// counter and error word live at chain[0..1] (cleared by the launch sequence itself); each
// of its waits gives up after chain_wait ticks and then sets chain[1].
static constexpr uint64_t CHAIN_WAIT_TICKS = 1ull << 27;  // ~1.3 s at the 100 MHz wall clock
// K: the transform kernels (K.chain must be set when chain is); ops / index: programs built
// for K.wnaf.
hipError_t launch_batch(const Launch &L, const DevTables *T, const uint32_t *ops, const uint32_t *index,
                        const StageKernels &K, hipStream_t s, LaunchHook hook, uint32_t *chain = nullptr,
                        uint64_t chain_wait = CHAIN_WAIT_TICKS);

// The drop-in entry points on the same column-batched transform kernels, for `rows`
// transforms of n = 2^logn points (n2 = n for the FK20 kinds), with D = rows * n * 144
// bytes of workspace (and, for PLAIN_POLY2H_FFT, tcT = rows * 16 * n * 32 bytes and
// digits = msm_plain_digit_bytes(rows, n)):
//   PLAIN_G1_FFT          out[r] = DFT(in[r])                         (g1p_fft)
//   PLAIN_G1_IFT          out[r] = DFT^-1(in[r]), scaled by 1/n       (g1p_ift)
//   PLAIN_HEXT_FFT2H_FFT  IFFT, zero upper half, FFT                  (fk20_hext_fft2h_fft)
//   PLAIN_POLY2H_FFT      Toeplitz + NTT (scaled by 1/n2), MSM over xpk (16 x n2 packed
//                         Montgomery bases), then as above            (fk20_poly2h_fft)
// out may alias in: every input is read into D before the first output is written.
enum PlainKind { PLAIN_G1_FFT = 0, PLAIN_G1_IFT = 1, PLAIN_HEXT_FFT2H_FFT = 2, PLAIN_POLY2H_FFT = 3 };
// flags: `rows` words of workspace (one word for PLAIN_POLY2H_FFT, set by the caller with
// launch_subgroup_groups over the bases).  Rows holding a point outside G1 are recomputed in
// the reference's order of operations (launch_ref_rows).
hipError_t launch_plain(PlainKind kind, uint32_t *out, const uint32_t *in, const uint32_t *xpk, unsigned rows,
                        unsigned logn, uint32_t *D, uint32_t *tcT, int8_t *digits, uint32_t *flags,
                        const DevTables *T, const Programs &P, hipStream_t s);

}  // namespace fk20
