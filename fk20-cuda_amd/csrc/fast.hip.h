// Batched FK20 path: twiddle-program encoding and launchers (fast_kernels.hip, ctx.cpp).
#pragma once
#include "common.hip.h"

namespace fk20 {

// A twiddle multiplication x <- w^idx * x is a straight-line program over the wNAF digits of
// the GLV split w^idx = k1 + k2 * lambda (k1, k2 < 2^128, phi(P) = lambda P), one 32-bit op
// per non-zero digit, read wave-uniformly:
//   bits 0..7   doublings to perform before this op (for the END op: trailing doublings)
//   bit  8      apply phi (the digit belongs to k2)
//   bits 9..11  table entry e: the odd multiple (2e + 1) P
//   bit  12     negate
//   bit  31     END
static constexpr uint32_t TWOP_DBL_MASK = 0xffu;
static constexpr uint32_t TWOP_PHI = 1u << 8;
static constexpr unsigned TWOP_ENTRY_SHIFT = 9;
static constexpr uint32_t TWOP_NEG = 1u << 12;
static constexpr uint32_t TWOP_END = 1u << 31;

// wNAF width used by the G1 transform kernels (odd multiples P, 3P, 5P, 7P)
static constexpr int TW_WNAF = 4;
// direction rows of the program index: index[dir * 512 + idx] = offset of w^(+-idx)'s program
static constexpr int TW_FWD = 0;
static constexpr int TW_INV = 1;

hipError_t launch_scale_bases(uint32_t *xs, uint32_t *inf, const uint32_t *xext, size_t count,
                              const uint32_t *scale, hipStream_t s);
hipError_t launch_table_build(uint32_t *table, uint32_t *zt, const uint32_t *xs, const uint32_t *inf,
                              unsigned base0, unsigned nb, unsigned win_bits, unsigned nwin, hipStream_t s);
hipError_t launch_toeplitz_fft_T(uint32_t *tcT, const uint32_t *poly, unsigned rows, unsigned N,
                                 const DevTables *T, hipStream_t s);
hipError_t launch_msm_comb(uint32_t *D, const uint32_t *tcT, const uint32_t *table, const uint32_t *inf,
                           unsigned rows, unsigned n2, unsigned win_bits, hipStream_t s);
// Kernel kinds of the batched path, reported by the per-launch timing of the context.
enum LaunchKind { LK_TOEPLITZ_FFT = 0, LK_MSM = 1, LK_DIT = 2, LK_MIDDLE = 3, LK_DIF = 4, LK_OUTPUT = 5 };
// Called after each launch when the context times launches (nullptr otherwise).
struct LaunchHook {
    void (*fn)(void *arg, int kind);
    void *arg;
    void operator()(int kind) const {
        if (fn) fn(arg, kind);
    }
};
hipError_t launch_g1_transforms(uint32_t *D, unsigned rows, unsigned n2, const uint32_t *ops,
                                const uint32_t *index, hipStream_t s, LaunchHook hook);
hipError_t launch_fft_output(uint32_t *h_fft, const uint32_t *D, unsigned rows, unsigned n2, hipStream_t s);

}  // namespace fk20
