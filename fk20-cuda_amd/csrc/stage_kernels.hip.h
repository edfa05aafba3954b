// The G1 transform kernels of the column-batched path, as templates over a schedule variant
// (wNAF width W, radix-2^2 or radix-2 butterflies R22, waves per SIMD OCC).  Each variant is
// instantiated in its own translation unit (stage_variant.hip) and published as a
// StageKernels table; a context picks its variant at run time (fk20_ctx_set_schedule).
#pragma once
#include "kutil.hip.h"

namespace fk20 {

// Per item with L = log2(n2), stage schedule over the 2*Lmax - 1 launches of a batch:
//   IFFT DIT stage s (0 <= s <= L-2)  at launch Lmax-1-(L-1)+s   (right-aligned)
//   middle                            at launch Lmax-1            (shared)
//   FFT DIF stage s (1 <= s <= L-1)   at launch Lmax-1+s          (left-aligned)
// so every item of a launch runs the same kind; it.stage holds s (or -1: idle).  Within an
// item, local stage u = 0 .. 2L-2 numbers its own stages: DIT s = u (u < L-1), the middle
// (u = L-1), DIF s = u - (L-1).  Butterfly b of a stage reads and writes two positions
// (in place), given by stage_positions().

// x <- w^(+-e) x, the sign by direction (index[dir * 512 + e] is the program of w^e for TW_FWD,
// of w^-e for TW_INV): w^+-128 by tw_mul_z3, w^256 = -1 by a negation, any other exponent by
// its program.  The exponent is wave-uniform, so is the branch.
template <int W>
__device__ __forceinline__ void twiddle(g1p &x, const uint32_t *ops, const uint32_t *index, unsigned dir,
                                        unsigned e) {
    const unsigned E = (dir == TW_FWD ? e : 512u - e) & 511u;  // x <- w^E x
    if (E == 128 || E == 384) tw_mul_z3(x, E == 128);
    else if (E == 256) fe_neg<4>(x.y, x.y);  // stored points have y < 4p
    else if (E) tw_mul<W>(x, ops + index[dir * 512 + e]);
}

// The butterflies' twiddle: in the radix-2^2 schedule always w^+-128 (idx 128), else any.
template <int W, bool R22>
__device__ __forceinline__ void bf_twiddle(g1p &x, const uint32_t *ops, const uint32_t *index, unsigned dir,
                                           unsigned idx) {
    if constexpr (R22) tw_mul_z3(x, dir == TW_FWD);
    else twiddle<W>(x, ops, index, dir, idx);
}

// IFFT decimation in time (bit-reversed input): x = D[p1] w^-idx; D[p0], D[p1] = y + x, y - x.
template <int W, bool R22>
__device__ __forceinline__ void dit_butterfly(uint32_t *D, unsigned rows, unsigned row, const Butterfly &f,
                                              const uint32_t *ops, const uint32_t *index) {
    g1p x, y, a;
    pt_load(x, D + ((size_t)f.p1 * rows + row) * 36);
    if (f.idx) bf_twiddle<W, R22>(x, ops, index, TW_INV, f.idx);
    pt_load(y, D + ((size_t)f.p0 * rows + row) * 36);
    g1_add(a, y, x);
    pt_store(D + ((size_t)f.p0 * rows + row) * 36, a);  // stored first: not live in the next add
    g1_neg(x, x);
    g1_add(x, y, x);
    pt_store(D + ((size_t)f.p1 * rows + row) * 36, x);
}

// Last IFFT stage, upper half dropped (h = first half), then the first FFT DIF stage whose
// upper input half is zero:  x = E_j + w^-idx O_j;  D[j] = x;  D[j + n2/2] = w^idx x.
template <int W, bool R22>
__device__ __forceinline__ void middle_butterfly(uint32_t *D, unsigned rows, unsigned row, const Butterfly &f,
                                                 const uint32_t *ops, const uint32_t *index) {
    g1p o, e;
    pt_load(o, D + ((size_t)f.p1 * rows + row) * 36);
    if (f.idx) bf_twiddle<W, R22>(o, ops, index, TW_INV, f.idx);
    pt_load(e, D + ((size_t)f.p0 * rows + row) * 36);
    g1_add(e, e, o);
    pt_store(D + ((size_t)f.p0 * rows + row) * 36, e);
    if (f.idx) bf_twiddle<W, R22>(e, ops, index, TW_FWD, f.idx);
    pt_store(D + ((size_t)f.p1 * rows + row) * 36, e);
}

// FFT decimation in frequency: D[p0], D[p1] = u + v, (u - v) w^idx.  In the radix-2^2
// schedule the w^128 of a quarter stage is applied by its own launch after the butterflies
// (k_quarter_multi): in the butterfly kernel those ladders ran at 0.9 ms per launch at 512
// rows where the same ladders in a separate launch take 0.5 ms.
template <int W, bool R22>
__device__ __forceinline__ void dif_butterfly(uint32_t *D, unsigned rows, unsigned row, const Butterfly &f,
                                              const uint32_t *ops, const uint32_t *index) {
    g1p u, v, a;
    pt_load(u, D + ((size_t)f.p0 * rows + row) * 36);
    pt_load(v, D + ((size_t)f.p1 * rows + row) * 36);
    g1_add(a, u, v);
    pt_store(D + ((size_t)f.p0 * rows + row) * 36, a);  // stored first: not live across tw_mul
    g1_neg(v, v);
    g1_add(v, u, v);
    if (!R22 && f.idx) bf_twiddle<W, R22>(v, ops, index, TW_FWD, f.idx);
    pt_store(D + ((size_t)f.p1 * rows + row) * 36, v);
}

// Runs local stage u's butterfly `local` of item `it` for this lane's row.  Lanes past the
// item's last row repeat the last row's work (same loads, same stores of the same values),
// so the wave never diverges: the persistent chain below relies on every lane staying
// converged around its wave-uniform steps.
template <int W>
__device__ __forceinline__ void run_butterfly(const Item &it, unsigned u, unsigned local, const uint32_t *ops,
                                              const uint32_t *index) {
    const unsigned rows = it.rows, rg = (rows + 63) / 64;
    const unsigned row = min((local % rg) * 64 + threadIdx.x, rows - 1);
    const Butterfly f = stage_positions<false>(u, it.logn2, local / rg);
    if (u + 1 < it.logn2) dit_butterfly<W, false>(it.D, rows, row, f, ops, index);
    else if (u + 1 == it.logn2) middle_butterfly<W, false>(it.D, rows, row, f, ops, index);
    else dif_butterfly<W, false>(it.D, rows, row, f, ops, index);
}

// One launch per stage: it.stage = s of this launch's kind (DIT/DIF), 0 for the middle.
// Each kernel inlines only its own butterfly (the chain kernel below needs all three).
// W: wNAF width of the twiddle programs; R22: radix-2^2 butterfly positions and twiddles
// (else radix 2, every butterfly carrying its full twiddle); OCC: waves per SIMD the
// register allocation must allow.
// Lanes past the item's last row repeat the last row's work (identical loads and stores
// within one wave) rather than leaving the wave: a ladder run by a wave with one active
// lane measured 2.3x slower than the same ladder with all 64 lanes active.
#define FK20_STAGE_KERNEL(NAME, BODY, USTAGE)                                                       \
    template <int W, bool R22, int OCC>                                                          \
    __global__ void __launch_bounds__(64, OCC) NAME(Launch L, const uint32_t *ops, const uint32_t *index) { \
        unsigned local;                                                                          \
        const Item &it = L.item[find_item(L, blockIdx.x, local)];                                \
        const unsigned rg = (it.rows + 63) / 64;                                                 \
        const unsigned row = min((local % rg) * 64 + threadIdx.x, it.rows - 1);                  \
        const Butterfly f = stage_positions<R22>((USTAGE), it.logn2, local / rg);                \
        BODY<W, R22>(it.D, it.rows, row, f, ops, index);                                         \
    }
FK20_STAGE_KERNEL(k_dit_multi, dit_butterfly, (unsigned)it.stage)
FK20_STAGE_KERNEL(k_middle_multi, middle_butterfly, it.logn2 - 1)
FK20_STAGE_KERNEL(k_dif_multi, dif_butterfly, it.logn2 - 1 + (unsigned)it.stage)
#undef FK20_STAGE_KERNEL

// The point-wise twiddle layer of odd DIF stage it.stage (radix-2^2 schedule): after that
// stage's butterflies in the FFT (dir = TW_FWD), before the butterflies of the transposed DIT
// stage in the IFFT (dir = TW_INV).  One wave per (nontrivial position, 64-row group), the
// positions taken in the cost order of the index's layer table.
template <int W, bool R22, int OCC>
__global__ void __launch_bounds__(64, OCC) k_twiddle_multi(Launch L, const uint32_t *ops, const uint32_t *index,
                                                          unsigned dir) {
    unsigned local;
    const Item &it = L.item[find_item(L, blockIdx.x, local)];
    const unsigned rg = (it.rows + 63) / 64;
    const unsigned row = min((local % rg) * 64 + threadIdx.x, it.rows - 1);
    const uint32_t item = index[TW_LAYER + layer_table_offset(it.logn2, (unsigned)it.stage, dir) + local / rg];
    uint32_t *p = it.D + ((size_t)(item & 0xffffu) * it.rows + row) * 36;
    g1p x;
    pt_load(x, p);
    twiddle<W>(x, ops, index, dir, item >> 16);
    pt_store(p, x);
}

// The w^128 twiddles of radix-2^2 DIF quarter stage it.stage (>= 1; stage 0 is inside the
// middle), after its butterflies: one wave per (twiddled butterfly, 64-row group), the
// butterflies of the launch's second half, whose difference output D[p1] takes w^128.
template <int W, bool R22, int OCC>
__global__ void __launch_bounds__(64, OCC) k_quarter_multi(Launch L) {
    unsigned local;
    const Item &it = L.item[find_item(L, blockIdx.x, local)];
    const unsigned rg = (it.rows + 63) / 64;
    const unsigned row = min((local % rg) * 64 + threadIdx.x, it.rows - 1);
    const unsigned b = (1u << (it.logn2 - 2)) + local / rg;
    const Butterfly f = stage_positions<true>(it.logn2 - 1 + (unsigned)it.stage, it.logn2, b);
    uint32_t *p = it.D + ((size_t)f.p1 * it.rows + row) * 36;
    g1p x;
    pt_load(x, p);
    tw_mul_z3(x, true);
    pt_store(p, x);
}

// ---------------------------------------------------------------- persistent transform chain
//
// All 2*Lmax - 1 stages in one launch of resident waves.  Work items (one butterfly of one
// 64-row group) are numbered stage by stage; a wave takes the next number from a counter,
// waits until the two positions it reads carry the flag of the item's previous local stage,
// runs the butterfly and flags both positions with u + 1.  A wave that finishes a cheap
// (untwiddled) butterfly moves on to later stages at once instead of idling at a launch
// boundary.  Items are taken in stage order, so every item an item depends on was taken
// earlier by a running wave: the waits always end.  Each wait is bounded all the same, by
// the constant-rate wall clock (C.wait_ticks, by default about 1.3 s at the 100 MHz
// reference; a legitimate wait is at most one pass, tens of ms); on expiry the wave records
// an error in ctl[1] and proceeds, and the host turns that word into an error status.

__device__ __forceinline__ unsigned flag_load(const uint32_t *f) {
    return __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int W, bool R22, int OCC>
__global__ void __launch_bounds__(64, OCC) k_chain(Chain C, const uint32_t *ops, const uint32_t *index) {
    // No lane-divergent branch anywhere in this loop: the work counter is bumped by lane 0's
    // addend (every lane executes the atomic, the others add 0) and every lane stores the
    // flags, so the wave-uniform values below always come from a fully converged wave.  A
    // wave can take at most `total` items, which bounds the loop whatever happens.
#pragma unroll 1
    for (unsigned iter = 0; iter < C.total; iter++) {
        const unsigned got =
            __hip_atomic_fetch_add(C.ctl, threadIdx.x == 0 ? 1u : 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned w = __builtin_amdgcn_readfirstlane(got);  // lane 0's: the value before its add
        if (w >= C.total) break;
        unsigned t = 0;
#pragma unroll 1
        while (w >= C.t_end[t]) t++;
        unsigned local = w - (t ? C.t_end[t - 1] : 0);
        unsigned k = 0;
#pragma unroll 1
        while (local >= C.blk_end[t][k]) k++;
        if (k) local -= C.blk_end[t][k - 1];
        const Item &it = C.L.item[k];
        const unsigned u = (unsigned)C.ustage[t][k];
        const unsigned rg = (it.rows + 63) / 64, g = local % rg;
        const Butterfly f = stage_positions<false>(u, it.logn2, local / rg);
        uint32_t *fl = it.flags + (size_t)g * (1u << it.logn2);
        if (u) {
            const uint64_t t0 = wall_clock64();
            bool expired = false;
#pragma unroll 1
            while (flag_load(fl + f.p0) < u || flag_load(fl + f.p1) < u) {
                if (wall_clock64() - t0 >= C.wait_ticks) {
                    expired = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
            if (expired) __hip_atomic_store(C.ctl + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        run_butterfly<W>(it, u, local, ops, index);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(fl + f.p0, u + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(fl + f.p1, u + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------- drop-in entry points

// D[pos][row] = in[row][j] (standard form -> packed Montgomery), pos = bitrev(j) when brev,
// times the scalar of the program at scale_ops when given.
template <int W, bool R22, int OCC>
__global__ void __launch_bounds__(64, OCC) k_plain_load(uint32_t *D, const uint32_t *in, unsigned rows,
                                                       unsigned logn, unsigned brev, const uint32_t *scale_ops) {
    const unsigned rg = (rows + 63) / 64, j = blockIdx.x / rg;
    const unsigned row = min((blockIdx.x % rg) * 64 + threadIdx.x, rows - 1);  // spare lanes repeat the last row
    uint32_t w[36];
    ld_words(w, in + ((size_t)row * (1u << logn) + j) * G1P_WORDS, 9);
    g1p p;
    g1_load_std(p, w);
    if (scale_ops) tw_mul<W>(p, scale_ops);
    pt_store(D + ((size_t)(brev ? bitrev(j, logn) : j) * rows + row) * 36, p);
}

// The kernel table of variant (W, R22, OCC); the persistent chain (radix 2 whatever the
// variant's butterflies) is published by the radix-2 variants only.
template <int W, bool R22, int OCC>
constexpr StageKernels make_stage_kernels() {
    StageKernels K{};
    K.dit = k_dit_multi<W, R22, OCC>;
    K.middle = k_middle_multi<W, R22, OCC>;
    K.dif = k_dif_multi<W, R22, OCC>;
    K.twiddle = k_twiddle_multi<W, R22, OCC>;
    if constexpr (!R22) K.chain = k_chain<W, R22, OCC>;
    if constexpr (R22) K.quarter = k_quarter_multi<W, R22, OCC>;
    K.plain_load = k_plain_load<W, R22, OCC>;
    K.wnaf = W;
    K.r22 = R22;
    K.occupancy = OCC;
    return K;
}

}  // namespace fk20
