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
