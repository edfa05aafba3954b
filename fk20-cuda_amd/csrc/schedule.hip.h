// Index maps of the column-batched G1 transforms: which positions each butterfly of each
// stage reads and writes, which twiddle it carries, and the point-wise twiddle layers of the
// radix-2^2 schedule.  Shared by the HIP kernels and, compiled by g++, by the CPU tests that
// run the schedule over Fr and check it against the DFT.
#pragma once
#include <stdint.h>

#ifndef FK20_HD
#define FK20_HD __host__ __device__ __forceinline__
#endif

namespace fk20 {

struct Butterfly {
    unsigned p0, p1;  // positions read and written
    unsigned idx;     // twiddle exponent (w^+-idx), 0: none
};

// R22 = false: radix 2, every butterfly carries its stage's twiddle (the persistent chain).
// R22 = true: radix 2^2 (the per-stage launches).  With the FFT's DIF stages numbered
// s = 0 .. L-1 (the IFFT's DIT stage u is the transpose of DIF stage L-1-u), stages pair up
// as (0, 1), (2, 3), ..: the even stage of a pair keeps only the w^(n/4) part (w^+-128 here,
// a 64-bit GLV program) of its twiddle, on the half of its butterflies whose offset in the
// block is at least half the span, and defers the rest to the odd stage, whose butterflies
// carry no twiddle; a separate point-wise layer (k_twiddle_multi, layer_twiddle) then applies
// the deferred and own twiddles of the odd stage, three per four points of each block.  A
// lone last DIF stage (L odd) has only trivial twiddles.  Both schedules compute the same
// transform exactly.
// A radix-2^2 stage that carries w^+-128 (on the butterflies whose offset j in a block of
// span 2^lh is at least 2^(lh-1)) numbers its butterflies as the middle does: the untwiddled
// ones first, the twiddled ones in the second half of the launch (b -> block and offset with
// the offset's top bit taken from b's top bit).  Measured on MI355X at 512 rows: the IFFT's
// w^-128 launches 1.1 -> 0.67 ms (blocks in launch order gave some SIMDs two ladders and
// others none).  The FFT's w^128 ladders run in a launch of their own after the butterflies
// (k_quarter_multi, over the second half of this numbering).
FK20_HD unsigned quarter_order(unsigned b, unsigned logn2, unsigned lh, unsigned &j) {
    const unsigned c = b >> (logn2 - 2), r = b & ((1u << (logn2 - 2)) - 1);
    j = (r & ((1u << (lh - 1)) - 1)) + (c << (lh - 1));
    return r >> (lh - 1);  // block
}

template <bool R22>
FK20_HD Butterfly stage_positions(unsigned u, unsigned logn2, unsigned b) {
    Butterfly f;
    if (u + 1 < logn2) {  // IFFT DIT stage s = u: span 2^s
        const unsigned s = u;
        unsigned j = b & ((1u << s) - 1), blk = b >> s;
        if (R22) {
            const unsigned sd = logn2 - 1 - s;  // the transposed DIF stage
            const bool quarter = (sd & 1) == 0 && sd + 1 < logn2;
            if (quarter) blk = quarter_order(b, logn2, s, j);
            f.idx = (quarter && 2 * j >= (1u << s)) ? 128u : 0u;
        } else {
            f.idx = j << (8 - s);
        }
        f.p0 = (blk << (s + 1)) + j;
        f.p1 = f.p0 + (1u << s);
    } else if (u + 1 == logn2) {  // middle: positions j, j + n2/2
        f.p0 = b;
        f.p1 = b + (1u << (logn2 - 1));
        f.idx = R22 ? (4 * b >= (1u << logn2) ? 128u : 0u) : b << (9 - logn2);
    } else {  // FFT DIF stage s = u - (L-1): span 2^(L-s-1)
        const unsigned s = u + 1 - logn2, lh = logn2 - s - 1;
        unsigned j = b & ((1u << lh) - 1), blk = b >> lh;
        if (R22) {
            const bool quarter = (s & 1) == 0 && s + 1 < logn2;
            if (quarter) blk = quarter_order(b, logn2, lh, j);
            f.idx = (quarter && 2 * j >= (1u << lh)) ? 128u : 0u;
        } else {
            f.idx = j << (8 - lh);
        }
        f.p0 = (blk << (lh + 1)) + j;
        f.p1 = f.p0 + (1u << lh);
    }
    return f;
}

// The point-wise twiddles of odd DIF stage sd (span h = 2^(L-1-sd) >= 2) in the radix-2^2
// schedule: in every block of 4h positions, quarter c (0..3) at offset j1 (0 < j1 < h) is
// multiplied by w^(m_c j1 2^(sd-1)), m = (0, 2, 1, 3), in units of w of order 2^L (w512^(e)
// with e scaled by 2^(9-L)).  Nontrivial twiddles are numbered k = 0 .. layer_count - 1.
struct LayerTw {
    unsigned q, e;  // position, exponent of w512
};

FK20_HD unsigned layer_count(unsigned sd, unsigned logn2) {
    const unsigned h = 1u << (logn2 - 1 - sd);
    return h < 2 ? 0 : 3 * (h - 1) * ((1u << logn2) / (4 * h));
}

FK20_HD LayerTw layer_twiddle(unsigned sd, unsigned logn2, unsigned k) {
    const unsigned lh = logn2 - 1 - sd, h = 1u << lh, per = 3 * (h - 1);
    const unsigned blk = k / per, r = k - blk * per;
    const unsigned c = 1 + r / (h - 1), j1 = 1 + r % (h - 1);
    LayerTw t;
    t.q = (blk << (lh + 2)) + (c << lh) + j1;
    t.e = ((c == 1 ? 2u : c == 2 ? 1u : 3u) * j1) << (sd - 1 + 9 - logn2);
    return t;
}

}  // namespace fk20
