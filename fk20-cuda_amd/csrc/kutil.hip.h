// Device helpers shared by the kernel translation units: vector loads and stores of words,
// points and field elements in the packed Montgomery memory form, and the per-launch item
// lookup of the batched path.
#pragma once
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

// ---------------------------------------------------------------- multi-item decoding
//
// Every launch below covers a list of items (row groups, each with its own setup, N and
// buffers).  A block finds its item from the per-item block prefix sums of the launch.

__device__ __forceinline__ unsigned find_item(const Launch &L, unsigned blk, unsigned &local) {
    unsigned k = 0, start = 0;
#pragma unroll 1
    while (k + 1 < L.nitems && blk >= L.blk_end[k]) {
        start = L.blk_end[k];
        k++;
    }
    local = blk - start;
    return k;
}

}  // namespace fk20
