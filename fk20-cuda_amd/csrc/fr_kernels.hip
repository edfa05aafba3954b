// Fr kernels: Toeplitz coefficient layout, Fr NTT of size 128/256/512, fused Toeplitz+NTT,
// and element-wise Fr arithmetic.
//
// NTT: one transform per workgroup of n/2 threads (one radix-2 butterfly per thread and
// stage), data staged in LDS in Montgomery form, decimation in time from a bit-reversed
// load to a natural-order result -- the same natural-order DFT as fr_fft.cu:28-168
// (forward) and fr_fft.cu:172-300 (inverse, scaled by 1/n).
#include "common.hip.h"

namespace fk20 {

// ---------------------------------------------------------------- Toeplitz layout

// tc[row][i][j] from polynomial[row][N] (FK20Py fk20_multi.py:92-93; the CUDA layout of
// fk20_poly2toeplitz_coefficients.cu:40-56 for N = 4096):
//   tc[i][0] = poly[N-1-i], tc[i][1..k+1] = 0, tc[i][k+2+m] = poly[31-i+16m], m = 0..k-3.
__device__ __forceinline__ int toeplitz_src(unsigned i, unsigned j, unsigned N) {
    const unsigned k = N / 16;
    if (j == 0) return (int)(N - 1 - i);
    if (j <= k + 1) return -1;
    return (int)(31 - i + 16 * (j - (k + 2)));
}

__global__ void __launch_bounds__(256) k_toeplitz(uint4 *tc, const uint4 *poly, unsigned rows, unsigned N) {
    const unsigned n2 = N / 8;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)rows * 16 * n2) return;
    const unsigned j = (unsigned)(e % n2), i = (unsigned)((e / n2) % 16);
    const size_t row = e / (16 * (size_t)n2);
    const int src = toeplitz_src(i, j, N);
    uint4 lo = make_uint4(0, 0, 0, 0), hi = lo;
    if (src >= 0) {
        lo = poly[2 * ((size_t)row * N + src)];
        hi = poly[2 * ((size_t)row * N + src) + 1];
    }
    tc[2 * e] = lo;
    tc[2 * e + 1] = hi;
}

// ---------------------------------------------------------------- NTT core

__device__ __forceinline__ void fs_load(fs &a, const uint32_t *p) {
    const uint4 *q = reinterpret_cast<const uint4 *>(p);
    uint4 x = q[0], y = q[1];
    a.l[0] = x.x; a.l[1] = x.y; a.l[2] = x.z; a.l[3] = x.w;
    a.l[4] = y.x; a.l[5] = y.y; a.l[6] = y.z; a.l[7] = y.w;
}

__device__ __forceinline__ void fs_store(uint32_t *p, const fs &a) {
    uint4 *q = reinterpret_cast<uint4 *>(p);
    q[0] = make_uint4(a.l[0], a.l[1], a.l[2], a.l[3]);
    q[1] = make_uint4(a.l[4], a.l[5], a.l[6], a.l[7]);
}

// In-place radix-2 DIT over LDS holding n = 2^logn Montgomery values in bit-reversed
// order; leaves the natural-order transform.  Stage s uses w_{2^(s+1)}^j = w_512^(j << (8-s)).
template <bool INV>
__device__ __forceinline__ void ntt_dit(fs *lds, unsigned logn, const DevTables *T) {
    const unsigned t = threadIdx.x;
    for (unsigned s = 0; s < logn; s++) {
        const unsigned h = 1u << s;
        const unsigned j = t & (h - 1);
        const unsigned l = ((t >> s) << (s + 1)) + j, r = l + h;
        const unsigned idx = j << (8 - s);
        fs x = lds[r], y = lds[l], w;
        fs_load(w, T->tw_mont[INV ? 512 - idx : idx]);
        fs_mul(x, x, w);
        fs_add(lds[l], y, x);
        fs_sub(lds[r], y, x);
        __syncthreads();
    }
}

template <unsigned LOGN, bool INV>
__global__ void __launch_bounds__(256) k_fr_ntt(uint32_t *out, const uint32_t *in, const DevTables *T) {
    constexpr unsigned n = 1u << LOGN, half = n / 2;
    __shared__ fs lds[n];
    const unsigned t = threadIdx.x;
    in += (size_t)blockIdx.x * n * FR_WORDS;
    out += (size_t)blockIdx.x * n * FR_WORDS;
#pragma unroll
    for (unsigned e = t; e < n; e += half) {
        fs a;
        fs_load(a, in + (size_t)e * FR_WORDS);
        fs_to_mont(a, a);
        lds[bitrev(e, LOGN)] = a;
    }
    __syncthreads();
    ntt_dit<INV>(lds, LOGN, T);
    fs sc;
    if (INV) fs_load(sc, T->inv2k_mont[LOGN]);
#pragma unroll
    for (unsigned e = t; e < n; e += half) {
        fs a = lds[e];
        if (INV) fs_mul(a, a, sc);
        fs_from_mont(a, a);
        fs_store(out + (size_t)e * FR_WORDS, a);
    }
}

// Toeplitz coefficients of one (row, i) built straight into LDS, then the NTT.
template <unsigned LOGN>
__global__ void __launch_bounds__(256) k_toeplitz_fft(uint32_t *tc_fft, const uint32_t *poly, unsigned N,
                                                      const DevTables *T) {
    constexpr unsigned n = 1u << LOGN, half = n / 2;
    __shared__ fs lds[n];
    const unsigned t = threadIdx.x;
    const unsigned i = blockIdx.x % 16;
    const size_t row = blockIdx.x / 16;
    poly += row * N * FR_WORDS;
    tc_fft += (size_t)blockIdx.x * n * FR_WORDS;
#pragma unroll
    for (unsigned e = t; e < n; e += half) {
        const int src = toeplitz_src(i, e, N);
        fs a;
        fs_zero(a);
        if (src >= 0) {
            fs_load(a, poly + (size_t)src * FR_WORDS);
            fs_to_mont(a, a);
        }
        lds[bitrev(e, LOGN)] = a;
    }
    __syncthreads();
    ntt_dit<false>(lds, LOGN, T);
#pragma unroll
    for (unsigned e = t; e < n; e += half) {
        fs a = lds[e];
        fs_from_mont(a, a);
        fs_store(tc_fft + (size_t)e * FR_WORDS, a);
    }
}

// ---------------------------------------------------------------- element-wise Fr ops

__global__ void k_fr_op(int op, uint32_t *out, const uint32_t *a, const uint32_t *b, size_t count) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    fs x, y, z;
    fs_load(x, a + e * FR_WORDS);
    fs_to_mont(x, x);
    fs_zero(y);
    if (op <= 2 && b) {  // binary ops only: add, sub, mul
        fs_load(y, b + e * FR_WORDS);
        fs_to_mont(y, y);
    }
    switch (op) {
    case 0: fs_add(z, x, y); break;
    case 1: fs_sub(z, x, y); break;
    case 2: fs_mul(z, x, y); break;
    case 3: fs_mul(z, x, x); break;
    case 4: { fs zero; fs_zero(zero); fs_sub(z, zero, x); } break;
    case 5: {  // x^(r-2)
        fs acc;
        fs_set(acc, FR_ONE);
        for (int w = 7; w >= 0; w--) {
            const uint32_t ew = FR_RM2[w];
            for (int bit = 31; bit >= 0; bit--) {
                fs_mul(acc, acc, acc);
                if ((ew >> bit) & 1) fs_mul(acc, acc, x);
            }
        }
        z = acc;
    } break;
    default: z = x; break;
    }
    fs_from_mont(z, z);
    fs_store(out + e * FR_WORDS, z);
}

// ---------------------------------------------------------------- launchers

hipError_t launch_toeplitz(uint32_t *tc, const uint32_t *poly, unsigned rows, unsigned N, hipStream_t s) {
    const size_t total = (size_t)rows * 16 * (N / 8);
    if (!total) return hipSuccess;
    const unsigned blocks = (unsigned)((total + 255) / 256);
    hipLaunchKernelGGL(k_toeplitz, dim3(blocks), dim3(256), 0, s, reinterpret_cast<uint4 *>(tc),
                       reinterpret_cast<const uint4 *>(poly), rows, N);
    return hipGetLastError();
}

template <bool INV>
static hipError_t ntt_dispatch(uint32_t *out, const uint32_t *in, unsigned count, unsigned n,
                               const DevTables *T, hipStream_t s) {
    switch (n) {
    case 128: hipLaunchKernelGGL((k_fr_ntt<7, INV>), dim3(count), dim3(64), 0, s, out, in, T); break;
    case 256: hipLaunchKernelGGL((k_fr_ntt<8, INV>), dim3(count), dim3(128), 0, s, out, in, T); break;
    case 512: hipLaunchKernelGGL((k_fr_ntt<9, INV>), dim3(count), dim3(256), 0, s, out, in, T); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fr_ntt(uint32_t *out, const uint32_t *in, unsigned count, unsigned n, bool inverse,
                         const DevTables *T, hipStream_t s) {
    if (!count) return hipSuccess;
    return inverse ? ntt_dispatch<true>(out, in, count, n, T, s) : ntt_dispatch<false>(out, in, count, n, T, s);
}

hipError_t launch_toeplitz_fft(uint32_t *tc_fft, const uint32_t *poly, unsigned rows, unsigned N,
                               const DevTables *T, hipStream_t s) {
    if (!rows) return hipSuccess;
    const unsigned blocks = rows * 16;
    switch (N) {
    case 1024: hipLaunchKernelGGL(k_toeplitz_fft<7>, dim3(blocks), dim3(64), 0, s, tc_fft, poly, N, T); break;
    case 2048: hipLaunchKernelGGL(k_toeplitz_fft<8>, dim3(blocks), dim3(128), 0, s, tc_fft, poly, N, T); break;
    case 4096: hipLaunchKernelGGL(k_toeplitz_fft<9>, dim3(blocks), dim3(256), 0, s, tc_fft, poly, N, T); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fr_op(int op, uint32_t *out, const uint32_t *a, const uint32_t *b, size_t count,
                        hipStream_t s) {
    if (!count) return hipSuccess;
    hipLaunchKernelGGL(k_fr_op, dim3((unsigned)((count + 63) / 64)), dim3(64), 0, s, op, out, a, b, count);
    return hipGetLastError();
}

}  // namespace fk20
