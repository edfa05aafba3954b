"""Canonical work model of one FK20 row, used for roofline figures.

Counted in 32x32-bit multiply-adds (mul32) of a 12 x 32-bit CIOS Montgomery multiply:
  Fp mul = 2*12^2 + 12 = 300 mul32,  Fr mul = 2*8^2 + 8 = 136 mul32
  dbl = 8 Fp mul (6M + 2S), add = 12 Fp mul (RCB 2015/1060 alg. 7)
  scalar mult SM = 255 dbl + 128 add = 3576 Fp mul; G1 butterfly BF = SM + 2 add = 3600
With n2 = N/8 (FFT size):
  NTT  = 16 * (n2/2) * log2(n2) Fr mul
  MSM  = 16 * n2 * (SM + add) Fp mul
  FFT  = (n2/2) * log2(n2) * BF Fp mul
  IFFT = FFT + n2 * SM Fp mul
The figure is fixed by this convention whatever algorithm runs (a comb MSM or
endomorphism-split twiddles execute fewer), so it measures useful work, not issued work.
"""
import math

FP_MUL32 = 2 * 12 * 12 + 12
FR_MUL32 = 2 * 8 * 8 + 8
DBL, ADD = 8, 12
SM = 255 * DBL + 128 * ADD
BF = SM + 2 * ADD

# MI355X integer VALU issue peak: 256 CUs x 64 lanes x 2.4 GHz, one v_mad_u64_u32 per lane
# per clock (the same issue rate as the published 78.6 TFLOP/s FP64 vector FMA figure).
PEAK_MUL32_PER_S = 256 * 64 * 2.4e9


def stage_fp_muls(N):
    n2 = N // 8
    lg = int(math.log2(n2))
    fft = (n2 // 2) * lg * BF
    return {
        "ntt_fr_muls": 16 * (n2 // 2) * lg,
        "msm": 16 * n2 * (SM + ADD),
        "ifft": fft + n2 * SM,
        "fft": fft,
    }


def row_mul32(N):
    s = stage_fp_muls(N)
    return (s["msm"] + s["ifft"] + s["fft"]) * FP_MUL32 + s["ntt_fr_muls"] * FR_MUL32


def stage_mul32(N):
    s = stage_fp_muls(N)
    return {
        "ntt": s["ntt_fr_muls"] * FR_MUL32,
        "msm": s["msm"] * FP_MUL32,
        "ifft": s["ifft"] * FP_MUL32,
        "fft": s["fft"] * FP_MUL32,
    }


# HBM bytes a row must move at minimum: polynomial in, h_fft out (projective)
def row_hbm_bytes(N):
    return N * 32 + (N // 8) * 144


def launch_mul32(N, kind, rows):
    """Canonical work of one launch of the batched path over `rows` rows.

    k_msm_comb: the MSM; k_transform_chain: all 2 log2(n2) radix-2 stages of (n2/2)
    butterflies (BF each).  The per-stage launches run the radix-2^2 schedule, which moves
    most twiddles out of the butterflies: their canonical share is split as
    (n2/2) * 2 add per butterfly launch (k_dit_stage / k_dif_stage; k_middle: the two
    stages it fuses) and one canonical scalar multiplication SM per twiddle a launch applies
    (k_twiddle_layer: the mean over a pass's layer launches; the butterfly launches' w^+-128
    twiddles likewise).  The IFFT's n2 scaling multiplications (folded into the setup tables
    here) are not attributed to any launch.
    """
    n2 = N // 8
    lg = int(math.log2(n2))
    if kind in ("k_dit_stage", "k_dif_stage", "k_middle"):
        if kind == "k_middle":
            tw = 2 * sum(1 for b in range(n2 // 2) if r22_butterfly_idx(lg - 1, lg, b))
            return rows * (2 * (n2 // 2) * 2 * ADD + tw * SM) * FP_MUL32
        us = range(lg - 1) if kind == "k_dit_stage" else range(lg, 2 * lg - 1)
        tw = sum(1 for u in us for b in range(n2 // 2) if r22_butterfly_idx(u, lg, b)) / len(us)
        return rows * ((n2 // 2) * 2 * ADD + tw * SM) * FP_MUL32
    if kind == "k_twiddle_layer":
        layers = r22_layers(N)
        tw = sum(len(r22_layer(sd, lg)) for sd, _ in layers) / len(layers)
        return rows * tw * SM * FP_MUL32
    stage = (n2 // 2) * BF * FP_MUL32
    return rows * {
        "k_msm_comb": 16 * n2 * (SM + ADD) * FP_MUL32,
        "k_transform_chain": 2 * lg * stage,
        "k_toeplitz_fft_T": 16 * (n2 // 2) * lg * FR_MUL32,
        "k_fft_output": 0,
    }[kind]


# ---------------------------------------------------------------- executed-work model
#
# Fp multiplications the batched path's algorithms actually perform (same 300-mul32
# convention; a square and each product of a shared-reduction pair count as one):
#   twiddle w^e: the library's own program (fk20_twiddle_program: GLV split, wNAF-5 digits)
#     run by the co-Z Jacobian ladder -- table P, 3P, .., 15P with one common Z (conversion 3,
#     doubling with update 7, seven co-Z additions 6 each, rescaling 34, table Z 1), one
#     Jacobian doubling (7) per doubling op, one mixed addition (11) per digit after the first,
#     one Fp mul per phi-digit, conversion back 4 and the infinity test 1
#   twiddle w^+-128: tw_mul_z3 -- conversion in and the infinity test 4, 63 Jacobian doublings
#     (7) and 5 mixed additions (11) of [|z|] x, conversion out with phi^2 5; w^256: a negation
#   butterfly: twiddle (if e != 0) + 2 complete additions (12 each)
#   MSM: one mixed addition (11 Fp mul) per non-zero signed window digit, 16 terms per column

R_MOD = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
MIXED_ADD = 11
J_DBL, J_MADD = 7, 11
WNAF_W = 5  # the library's default width (fk20_twiddle_selfcheck); a test ties the two


def ladder_fixed(wnaf=WNAF_W):
    """Fp muls of a ladder outside its doublings and additions: conversion in, doubling with
    update, co-Z additions, rescaling and the table's Z, conversion out, infinity test."""
    nt = 1 << (wnaf - 2)
    return 3 + 7 + 6 * (nt - 1) + 4 * (nt - 1) + (nt - 2) + 1 + 4 + 1


LADDER_FIXED = ladder_fixed()

TWOP_DBL_MASK, TWOP_PHI, TWOP_PHI2, TWOP_END = 0xFF, 1 << 8, 1 << 13, 1 << 31

_programs = {}


def program_counts(ops):
    """(doublings, additions, phi_muls, ops) of a program in the library's op encoding."""
    ops = [int(x) for x in ops]
    dbl = sum(op & TWOP_DBL_MASK for op in ops)
    digits = [op for op in ops if not op & TWOP_END]
    phi = sum(1 for op in digits if op & (TWOP_PHI | TWOP_PHI2))
    return dbl, max(len(digits) - 1, 0), phi, len(ops)


def twiddle_program(e_idx, inverse=False, wnaf=WNAF_W):
    """(doublings, additions, phi_muls, ops) of the library's program for w^e (w of order 512)."""
    key = (1 if inverse else 0, e_idx, wnaf)
    if key not in _programs:
        from . import prover
        _programs[key] = program_counts(prover.twiddle_program(*key))
    return _programs[key]


Z3_LADDER = 4 + 63 * J_DBL + 5 * J_MADD + 5


def twiddle_fp_muls(e_idx, inverse=False, wnaf=WNAF_W):
    """Executed Fp muls of the transform kernels' multiplication by w^(+-e_idx)
    (stage_kernels.hip.h twiddle): the short forms of w^+-128 and w^256, else the program."""
    e = (-e_idx if inverse else e_idx) % 512
    if e in (128, 384):
        return Z3_LADDER
    if e in (0, 256):
        return 0
    dbl, adds, phi, _ = twiddle_program(e_idx, inverse, wnaf)
    return ladder_fixed(wnaf) + dbl * J_DBL + adds * J_MADD + phi


def _quarter_j(b, lg, lh):
    """Offset j of launch index b in a renumbered w^+-128 stage of span 2^lh
    (schedule.hip.h quarter_order: the untwiddled half of the launch first)."""
    c, r = b >> (lg - 2), b & ((1 << (lg - 2)) - 1)
    return (r & ((1 << (lh - 1)) - 1)) + (c << (lh - 1))


def r22_butterfly_idx(u, lg, b):
    """Twiddle exponent (units of w512) of launch index b of local stage u in the radix-2^2
    schedule of the per-stage launches (schedule.hip.h stage_positions<true>): only the
    even DIF stages (and the transposed DIT stages) keep a w^(n/4) = w512^128 part."""
    n2 = 1 << lg
    if u + 1 < lg:  # DIT stage s = u
        s = u
        sd = lg - 1 - s
        if not (sd % 2 == 0 and sd + 1 < lg):
            return 0
        return 128 if 2 * _quarter_j(b, lg, s) >= (1 << s) else 0
    if u + 1 == lg:  # middle
        return 128 if 4 * b >= n2 else 0
    s = u + 1 - lg
    lh = lg - s - 1
    if not (s % 2 == 0 and s + 1 < lg):
        return 0
    return 128 if 2 * _quarter_j(b, lg, lh) >= (1 << lh) else 0


def r22_layer(sd, lg):
    """(position, exponent of w512) of every nontrivial twiddle of the point-wise layer of odd
    DIF stage sd (fast_kernels.hip layer_twiddle)."""
    h = 1 << (lg - 1 - sd)
    if h < 2:
        return []
    n2 = 1 << lg
    out = []
    for k in range(3 * (h - 1) * (n2 // (4 * h))):
        blk, r = divmod(k, 3 * (h - 1))
        c, j1 = 1 + r // (h - 1), 1 + r % (h - 1)
        out.append((blk * 4 * h + c * h + j1, ({1: 2, 2: 1, 3: 3}[c] * j1) << (sd - 1 + 9 - lg)))
    return out


def r22_layers(N):
    """The twiddle-layer launches of one pass, as (sd, inverse): before DIT stages (inverse)
    and after DIF stages (forward)."""
    lg = int(math.log2(N // 8))
    dit = [(lg - 1 - u, True) for u in range(lg - 1) if (lg - 1 - u) % 2 == 1 and r22_layer(lg - 1 - u, lg)]
    dif = [(s, False) for s in range(1, lg) if s % 2 == 1 and r22_layer(s, lg)]
    return dit + dif


def _radix2_idx(u, lg, b):
    """Twiddle exponent (units of w512) of butterfly b of local stage u in the radix-2
    schedule (schedule.hip.h stage_positions<false>)."""
    if u + 1 < lg:
        return (b & ((1 << u) - 1)) << (8 - u)
    if u + 1 == lg:
        return b << (9 - lg)
    lh = 2 * lg - 2 - u
    return (b & ((1 << lh) - 1)) << (8 - lh)


def executed_launch_fp_muls(N, kind, stage=None, window=8, wnaf=WNAF_W, r22=True):
    """Executed Fp muls of one launch of the batched path, per row.  k_dit_stage /
    k_dif_stage / k_middle / k_twiddle_layer (stage = (sd, inverse)) follow the schedule of
    the per-stage launches (radix-2^2 with twiddle layers, or radix 2 with r22 False);
    k_transform_chain the radix-2 schedule of the persistent chain.  wnaf: the programs'
    digit width."""
    n2 = N // 8
    lg = int(math.log2(n2))
    bidx = r22_butterfly_idx if r22 else _radix2_idx

    def tw(e, inverse=False):
        return twiddle_fp_muls(e, inverse, wnaf)

    if kind == "k_msm_comb":
        # ceil(256/w) signed windows per term; zero digits (probability ~2^-w) ignored
        return 16 * n2 * ((256 + window - 1) // window) * MIXED_ADD
    if kind in ("k_dit_stage", "k_dif_stage"):
        u = stage if kind == "k_dit_stage" else lg - 1 + stage
        total = 0
        for b in range(n2 // 2):
            idx = bidx(u, lg, b)
            total += 2 * ADD + (tw(idx, inverse=(kind == "k_dit_stage")) if idx else 0)
        return total
    if kind == "k_middle":
        total = 0
        for b in range(n2 // 2):
            idx = bidx(lg - 1, lg, b)
            total += ADD + (tw(idx, True) + tw(idx, False) if idx else 0)
        return total
    if kind == "k_twiddle_layer":
        sd, inverse = stage
        return sum(tw(e, inverse) for _, e in r22_layer(sd, lg))
    if kind == "k_transform_chain":
        total = 0
        for u in range(lg - 1):  # radix-2 DIT stages
            for b in range(n2 // 2):
                idx = _radix2_idx(u, lg, b)
                total += 2 * ADD + (tw(idx, True) if idx else 0)
        for j in range(n2 // 2):  # middle
            idx = _radix2_idx(lg - 1, lg, j)
            total += ADD + (tw(idx, True) + tw(idx, False) if idx else 0)
        for u in range(lg, 2 * lg - 1):  # radix-2 DIF stages
            for b in range(n2 // 2):
                idx = _radix2_idx(u, lg, b)
                total += 2 * ADD + (tw(idx) if idx else 0)
        return total
    raise KeyError(kind)


def executed_row_fp_muls(N, window=8, wnaf=WNAF_W, r22=True, chain=False):
    lg = int(math.log2(N // 8))
    t = executed_launch_fp_muls(N, "k_msm_comb", window=window)
    if chain:
        return t + executed_launch_fp_muls(N, "k_transform_chain", wnaf=wnaf)
    t += sum(executed_launch_fp_muls(N, "k_dit_stage", s, wnaf=wnaf, r22=r22) for s in range(lg - 1))
    t += executed_launch_fp_muls(N, "k_middle", wnaf=wnaf, r22=r22)
    t += sum(executed_launch_fp_muls(N, "k_dif_stage", s, wnaf=wnaf, r22=r22) for s in range(1, lg))
    if r22:
        t += sum(executed_launch_fp_muls(N, "k_twiddle_layer", st, wnaf=wnaf) for st in r22_layers(N))
    return t
