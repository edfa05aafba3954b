// Host build of the device arithmetic headers (fp.hip.h, fr.hip.h, g1.hip.h) for CPU tests:
// the same source the HIP kernels inline, compiled by g++ with FK20_HD = inline.
// Every function takes and returns the reference ABI layout (standard form, u64 limbs).
#define FK20_HD inline
#define __host__
#define __device__
#define FK20_UNIFORM(x) (x)
#include "../../fk20-cuda_amd/csrc/ladder.hip.h"
#include "../../fk20-cuda_amd/csrc/schedule.hip.h"

#include <cstring>

using namespace fk20;

static void ld(fe &a, const uint64_t *x) {
    uint32_t w[12];
    memcpy(w, x, 48);
    fe t;
    fe_unpack(t, w);
    fe_to_mont(a, t);
}
static void st(uint64_t *x, const fe &a) {
    fe t;
    fe_from_mont(t, a);
    uint32_t w[12];
    fe_pack(w, t);
    memcpy(x, w, 48);
}
static void ldp(g1p &p, const uint64_t *x) {
    ld(p.x, x);
    ld(p.y, x + 6);
    ld(p.z, x + 12);
}
static void stp(uint64_t *x, const g1p &p) {
    st(x, p.x);
    st(x + 6, p.y);
    st(x + 12, p.z);
}

extern "C" {
// op: 0 add, 1 sub, 2 mul, 3 sqr, 4 neg, 5 inv, 6 reduce; `reps` repeats z = op(z, b) to
// exercise lazy bounds (add/sub results feed the next step without reduction).
void host_fp_op(int op, uint64_t *z, const uint64_t *a, const uint64_t *b, int reps) {
    fe x, y, r;
    ld(x, a);
    if (b) ld(y, b);
    r = x;
    for (int i = 0; i < reps; i++) {
        switch (op) {
        case 0: fe_add(r, r, y); break;
        case 1: fe_sub<2>(r, r, y); break;
        case 2: fe_mul(r, r, y); break;
        case 3: fe_sqr(r, r); break;
        case 4: fe_neg<2>(r, r); break;
        case 5: fe_inv(r, r); break;
        default: break;
        }
    }
    st(z, r);
}

void host_fr_mul(uint64_t *z, const uint64_t *a, const uint64_t *b) {
    fs x, y;
    memcpy(x.l, a, 32);
    memcpy(y.l, b, 32);
    fs_to_mont(x, x);
    fs_to_mont(y, y);
    fs_mul(x, x, y);
    fs_from_mont(x, x);
    memcpy(z, x.l, 32);
}

void host_g1_add(uint64_t *r, const uint64_t *p, const uint64_t *q) {
    g1p a, b;
    ldp(a, p);
    ldp(b, q);
    g1_add(a, a, b);
    stp(r, a);
}

// Raw limbs in and out (14 x u32 each, Montgomery domain untouched): z = (ab + cd)/2^406
// mod p, as fe_mul2 computes it, so limb extremes can be driven directly.
void host_fp_mul2_raw(uint32_t *z, const uint32_t *a, const uint32_t *b, const uint32_t *c,
                      const uint32_t *d) {
    fe x, y, u, v, r;
    memcpy(x.l, a, sizeof x.l);
    memcpy(y.l, b, sizeof y.l);
    memcpy(u.l, c, sizeof u.l);
    memcpy(v.l, d, sizeof v.l);
    fe_mul2(r, x, y, u, v);
    memcpy(z, r.l, sizeof r.l);
}

void host_fp_mul_raw(uint32_t *z, const uint32_t *a, const uint32_t *b) {
    fe x, y, r;
    memcpy(x.l, a, sizeof x.l);
    memcpy(y.l, b, sizeof y.l);
    fe_mul(r, x, y);
    memcpy(z, r.l, sizeof r.l);
}

// r = p + (x, y) with the affine operand given as a g1p_t whose z is ignored.
void host_g1_add_aff(uint64_t *r, const uint64_t *p, const uint64_t *q) {
    g1p a;
    fe x, y;
    ldp(a, p);
    ld(x, q);
    ld(y, q + 6);
    g1_add_aff(a, a, x, y);
    stp(r, a);
}

// 48-byte compressed serialisation of a g1p_t.
void host_g1_compress(uint8_t *out, const uint64_t *p) {
    g1p a;
    ldp(a, p);
    uint32_t w[12];
    g1_compress(w, a);
    memcpy(out, w, 48);
}

void host_g1_dbl(uint64_t *r, const uint64_t *p) {
    g1p a;
    ldp(a, p);
    g1_dbl(a, a);
    stp(r, a);
}

void host_g1_mul(uint64_t *r, const uint64_t *p, const uint64_t *k) {
    g1p a;
    ldp(a, p);
    uint32_t kw[8];
    memcpy(kw, k, 32);
    g1_mul(a, a, kw);
    stp(r, a);
}

// Chain of mixed operations that keeps intermediate values in the lazy (unreduced) range.
void host_g1_chain(uint64_t *r, const uint64_t *p, int steps) {
    g1p a, b;
    ldp(a, p);
    b = a;
    for (int i = 0; i < steps; i++) {
        g1_dbl(a, a);
        g1_add(a, a, b);
        g1p n;
        g1_neg(n, b);
        g1_add(b, a, n);
    }
    stp(r, b);
}

// The transform kernels' fixed-scalar ladder (ladder.hip.h, table width FK20_TW_WNAF) on a
// g1p_t, running the program `ops` (as fk20_twiddle_program returns it).
void host_tw_mul(uint64_t *r, const uint64_t *p, const uint32_t *ops) {
    g1p a;
    ldp(a, p);
    tw_mul<TW_WNAF>(a, ops);
    stp(r, a);
}

// The same ladder at table width w (3..5), for programs built at that width.
void host_tw_mul_w(uint64_t *r, const uint64_t *p, const uint32_t *ops, int w) {
    g1p a;
    ldp(a, p);
    if (w == 3) tw_mul<3>(a, ops);
    else if (w == 4) tw_mul<4>(a, ops);
    else tw_mul<5>(a, ops);
    stp(r, a);
}

// The short form of the w^+-128 twiddles (ladder.hip.h): w128 = 1 for w^128, 0 for w^-128.
void host_tw_mul_z3(uint64_t *r, const uint64_t *p, int w128) {
    g1p a;
    ldp(a, p);
    tw_mul_z3(a, w128 != 0);
    stp(r, a);
}

// The kernels' membership test of the order-r subgroup (ladder.hip.h).
int host_g1_in_subgroup(const uint64_t *p) {
    g1p a;
    ldp(a, p);
    return g1_in_subgroup(a) ? 1 : 0;
}

// Jacobian doubling / mixed addition, g1p_t slots holding Jacobian (X, Y, Z) and the affine
// operand (x, y) in the first two coordinates; `reps` repeats the operation on its output.
void host_g1j_dbl(uint64_t *r, const uint64_t *p, int reps) {
    g1j a;
    ld(a.x, p);
    ld(a.y, p + 6);
    ld(a.z, p + 12);
    for (int i = 0; i < reps; i++) g1j_dbl(a, a);
    st(r, a.x);
    st(r + 6, a.y);
    st(r + 12, a.z);
}

void host_g1j_madd(uint64_t *r, const uint64_t *p, const uint64_t *q) {
    g1j a;
    fe x, y;
    ld(a.x, p);
    ld(a.y, p + 6);
    ld(a.z, p + 12);
    ld(x, q);
    ld(y, q + 6);
    g1j_madd(a, a, x, y);
    st(r, a.x);
    st(r + 6, a.y);
    st(r + 12, a.z);
}

// The transform kernels' index maps (schedule.hip.h): butterfly b of local stage u as
// (p0, p1, twiddle exponent) in the radix-2^2 (r22 = 1) or radix-2 schedule, and nontrivial
// twiddle k of the layer of odd DIF stage sd as (position, exponent of w512).
void host_stage_positions(int r22, unsigned u, unsigned logn2, unsigned b, unsigned *out) {
    const Butterfly f = r22 ? stage_positions<true>(u, logn2, b) : stage_positions<false>(u, logn2, b);
    out[0] = f.p0;
    out[1] = f.p1;
    out[2] = f.idx;
}

unsigned host_layer_count(unsigned sd, unsigned logn2) { return layer_count(sd, logn2); }

void host_layer_twiddle(unsigned sd, unsigned logn2, unsigned k, unsigned *out) {
    const LayerTw t = layer_twiddle(sd, logn2, k);
    out[0] = t.q;
    out[1] = t.e;
}
}
