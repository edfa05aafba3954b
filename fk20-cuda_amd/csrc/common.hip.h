// Internal interface between the HIP kernel translation units and the C-ABI host code.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "g1.hip.h"

namespace fk20 {

// Per-device constant tables, built once on the host (api.cpp) and copied to HBM.
struct DevTables {
    uint32_t tw_mont[513][8];     // w^i, w = 5^((r-1)/512), Montgomery form
    uint32_t tw_std[513][8];      // w^i, standard form (scalar-multiplier input)
    uint32_t inv2k_mont[10][8];   // 2^-j mod r, Montgomery form, j = 0..9
    uint32_t inv2k_std[10][8];    // 2^-j mod r, standard form
    uint32_t inv2k_w384_std[10][8];  // 2^-j w^384 = 2^-j w^-128 mod r, standard form
};

// Reference ABI element sizes in 32-bit words
static constexpr int FR_WORDS = 8;    // fr_t  (fr.cuh:24)
static constexpr int G1P_WORDS = 36;  // g1p_t (g1.cuh:27)
static constexpr int G1A_WORDS = 24;  // g1a_t (g1.cuh:20)

__host__ __device__ __forceinline__ unsigned bitrev(unsigned x, unsigned bits) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_bitreverse32(x) >> (32 - bits);
#else
    unsigned r = 0;
    for (unsigned i = 0; i < bits; i++) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
#endif
}

// Host copy of the constant tables (api.cpp).
void build_host_tables(DevTables &T);

// FK20_* status of a HIP error, recorded as the thread's last HIP error (api.cpp).
int status_from_hip(hipError_t e);

// Constant tables of the current device, built on first use (api.cpp).  Returns nullptr
// and sets *err on failure.
const DevTables *current_device_tables(hipError_t *err);

// ---- launchers (one per kernel family); all asynchronous on `s` ----
hipError_t launch_toeplitz(uint32_t *tc, const uint32_t *poly, unsigned rows, unsigned N, hipStream_t s);
hipError_t launch_toeplitz_fft(uint32_t *tc_fft, const uint32_t *poly, unsigned rows, unsigned N,
                               const DevTables *T, hipStream_t s);
hipError_t launch_fr_ntt(uint32_t *out, const uint32_t *in, unsigned count, unsigned n, bool inverse,
                         const DevTables *T, hipStream_t s);
hipError_t launch_hext2h(uint32_t *h, unsigned rows, unsigned n2, hipStream_t s);
hipError_t launch_std_to_packed(uint32_t *packed, const uint32_t *std_pts, size_t count, hipStream_t s);
// fk20_msm over per-call bases (16 x n2 packed Montgomery points); `digits` is workspace of
// msm_plain_digit_bytes(rows, n2) bytes (the scalars' Booth digits).
size_t msm_plain_digit_bytes(unsigned rows, unsigned n2);
hipError_t launch_msm_plain(uint32_t *hext_fft, const uint32_t *tc_fft, const uint32_t *xext_packed, int8_t *digits,
                            unsigned rows, unsigned n2, hipStream_t s);
// The same MSM from transposed scalars tcT[j][i][row] into D[bitrev(j)][row] (packed
// Montgomery), the layouts of the column-batched transforms.
hipError_t launch_msm_plain_T(uint32_t *D, const uint32_t *tcT, const uint32_t *xext_packed, int8_t *digits,
                              unsigned rows, unsigned n2, hipStream_t s);
// blocks x 256 lanes, each 8 x iters v_mad_u64_u32; out receives blocks * 256 words.
hipError_t launch_mad_probe(uint64_t *out, unsigned blocks, unsigned iters, hipStream_t s);
// blocks x 256 lanes, two chains of iters Fp multiplications each in the 14 x 29-bit (limbs 29)
// or 12 x 32-bit CIOS (limbs 32) layout; out receives blocks * 256 * 12 words.
hipError_t launch_fp_mul_probe(int limbs, uint32_t *out, const uint32_t *xs, const uint32_t *y, unsigned blocks,
                               unsigned iters, hipStream_t s);
hipError_t launch_setup_gather(uint32_t *xext, const uint32_t *srs, unsigned N, hipStream_t s);
hipError_t launch_to_affine(uint32_t *out, const uint32_t *in, size_t count, hipStream_t s);
hipError_t launch_compress(uint32_t *out, const uint32_t *in, size_t count, hipStream_t s);
hipError_t launch_fp_op(int op, uint32_t *out, const uint32_t *a, const uint32_t *b, size_t count,
                        hipStream_t s);
hipError_t launch_fr_op(int op, uint32_t *out, const uint32_t *a, const uint32_t *b, size_t count,
                        hipStream_t s);
hipError_t launch_g1_op(int op, uint32_t *out, const uint32_t *p, const uint32_t *q, const uint32_t *k,
                        size_t count, hipStream_t s);

}  // namespace fk20
