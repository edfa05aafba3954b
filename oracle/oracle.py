"""ctypes binding of the CPU oracle (libfk20_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check in __graft_entry__.py
and by the cpu_baseline leg of bench.py, always as the parity reference or the timed CPU baseline.

Arrays use the reference ABI (FK20-CUDA fr.cuh:24, g1.cuh:20-30) as numpy uint64:
  fr  -> shape (..., 4)    g1p -> shape (..., 18)    g1a -> shape (..., 12)
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libfk20_oracle.so")

R_MOD = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
P_MOD = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        _lib = ctypes.CDLL(LIBPATH)
        _lib.or_default_threads.restype = ctypes.c_int
        _lib.or_g1p_match_affine_n.restype = ctypes.c_int64
        for fn in ["or_fr_eq", "or_fp_eq", "or_g1p_is_inf", "or_g1p_is_point", "or_g1p_eq"]:
            getattr(_lib, fn).restype = ctypes.c_int
    return _lib


def _p(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


# ---- integer <-> limb helpers ----

def int_to_limbs(v, n):
    return np.array([(v >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(n)], dtype=np.uint64)


def limbs_to_int(a):
    v = 0
    for i, w in enumerate(np.asarray(a, dtype=np.uint64).tolist()):
        v |= int(w) << (64 * i)
    return v


def fr(v):
    return int_to_limbs(v, 4)


def fp(v):
    return int_to_limbs(v, 6)


def g1p_from_ints(x, y, z):
    return np.concatenate([fp(x), fp(y), fp(z)])


# ---- field ops ----

def _binop(fn, n, x, y):
    z = np.zeros(n, dtype=np.uint64)
    getattr(lib(), fn)(_p(z), _p(np.ascontiguousarray(x, dtype=np.uint64)),
                       _p(np.ascontiguousarray(y, dtype=np.uint64)))
    return z


def _unop(fn, n, x):
    z = np.zeros(n, dtype=np.uint64)
    getattr(lib(), fn)(_p(z), _p(np.ascontiguousarray(x, dtype=np.uint64)))
    return z


def fr_op(op, *args):
    return (_binop if len(args) == 2 else _unop)("or_fr_" + op, 4, *args)


def fp_op(op, *args):
    return (_binop if len(args) == 2 else _unop)("or_fp_" + op, 6, *args)


def fr_root_of_unity(n):
    z = np.zeros(4, dtype=np.uint64)
    lib().or_fr_root_of_unity(_p(z), ctypes.c_uint64(n))
    return z


def fr_roots515():
    z = np.zeros((515, 4), dtype=np.uint64)
    lib().or_fr_roots515(_p(z))
    return z


# ---- G1 ----

def g1p_inf():
    z = np.zeros(18, dtype=np.uint64)
    lib().or_g1p_inf(_p(z))
    return z


def g1p_gen():
    z = np.zeros(18, dtype=np.uint64)
    lib().or_g1p_gen(_p(z))
    return z


def g1p_op(op, *args):
    if op in ("add", "sub"):
        return _binop("or_g1p_" + op, 18, *args)
    if op == "add_affine":
        return _binop("or_g1p_add_affine", 18, *args)
    if op == "mul":
        return _binop("or_g1p_mul", 18, *args)
    return _unop("or_g1p_" + op, 18, *args)


def g1p_eq(p, q):
    return bool(lib().or_g1p_eq(_p(np.ascontiguousarray(p, dtype=np.uint64)),
                                _p(np.ascontiguousarray(q, dtype=np.uint64))))


def g1p_is_point(p):
    return bool(lib().or_g1p_is_point(_p(np.ascontiguousarray(p, dtype=np.uint64))))


def g1p_is_inf(p):
    return bool(lib().or_g1p_is_inf(_p(np.ascontiguousarray(p, dtype=np.uint64))))


def to_affine(p):
    p = np.ascontiguousarray(p, dtype=np.uint64)
    count = p.size // 18
    a = np.zeros(p.shape[:-1] + (12,), dtype=np.uint64)
    lib().or_g1p_to_affine_n(_p(a), _p(p), ctypes.c_int64(count))
    return a


# ---- transforms and FK20 ----

def fr_fft(x, inverse=False):
    x = np.ascontiguousarray(x, dtype=np.uint64)
    n = x.shape[0]
    out = np.zeros_like(x)
    lib().or_fr_fft(_p(out), _p(x), ctypes.c_int(n), ctypes.c_int(int(inverse)))
    return out


def g1p_fft(x, inverse=False):
    x = np.ascontiguousarray(x, dtype=np.uint64)
    n = x.shape[0]
    out = np.zeros_like(x)
    lib().or_g1p_fft(_p(out), _p(x), ctypes.c_int(n), ctypes.c_int(int(inverse)))
    return out


def poly2toeplitz_coefficients(poly):
    poly = np.ascontiguousarray(poly, dtype=np.uint64)
    N = poly.shape[0]
    tc = np.zeros((16, N // 8, 4), dtype=np.uint64)
    lib().or_fk20_poly2toeplitz_coefficients(_p(tc), _p(poly), ctypes.c_int(N))
    return tc


def msm(tc_fft, xext_fft):
    tc_fft = np.ascontiguousarray(tc_fft, dtype=np.uint64)
    xext_fft = np.ascontiguousarray(xext_fft, dtype=np.uint64)
    n2 = tc_fft.shape[1]
    out = np.zeros((n2, 18), dtype=np.uint64)
    lib().or_fk20_msm(_p(out), _p(tc_fft), _p(xext_fft), ctypes.c_int(n2))
    return out


def hext_fft2h_fft(hext_fft):
    hext_fft = np.ascontiguousarray(hext_fft, dtype=np.uint64)
    out = np.zeros_like(hext_fft)
    lib().or_fk20_hext_fft2h_fft(_p(out), _p(hext_fft), ctypes.c_int(hext_fft.shape[0]))
    return out


def poly2h_fft(poly, xext_fft, nthreads=0):
    """poly: (rows, N, 4) -> h_fft (rows, N/8, 18)."""
    poly = np.ascontiguousarray(poly, dtype=np.uint64)
    if poly.ndim == 2:
        poly = poly[None]
    rows, N = poly.shape[0], poly.shape[1]
    out = np.zeros((rows, N // 8, 18), dtype=np.uint64)
    lib().or_fk20_poly2h_fft(_p(out), _p(poly), _p(np.ascontiguousarray(xext_fft, dtype=np.uint64)),
                             ctypes.c_int(N), ctypes.c_int(rows), ctypes.c_int(nthreads))
    return out


def setup_srs(tau, count):
    out = np.zeros((count, 18), dtype=np.uint64)
    lib().or_setup_srs(_p(out), _p(fr(tau)), ctypes.c_int(count))
    return out


def commit(poly, srs):
    """KZG commitments sum_i poly[r][i] * srs[i] of rows (rows, N, 4) over srs (>= N, 18):
    FK20Py commit_to_poly (kzg_proofs.py:159), by direct double-and-add."""
    poly = np.ascontiguousarray(poly, dtype=np.uint64)
    rows, N = poly.shape[0], poly.shape[1]
    srs = np.ascontiguousarray(srs, dtype=np.uint64)
    out = np.zeros((rows, 18), dtype=np.uint64)
    lib().or_commit(_p(out), _p(poly), _p(srs), ctypes.c_int(N), ctypes.c_int(rows))
    return out


def commit_closed_form(poly_row, tau):
    """[f(tau)]G for one row: the commitment of a setup with secret tau."""
    f = 0
    for c in reversed([limbs_to_int(x) for x in np.asarray(poly_row)]):
        f = (f * tau + c) % R_MOD
    return g1p_op("mul", g1p_gen(), fr(f))


def setup2xext_fft(srs, N):
    out = np.zeros((16, N // 8, 18), dtype=np.uint64)
    lib().or_setup2xext_fft(_p(out), _p(np.ascontiguousarray(srs, dtype=np.uint64)), ctypes.c_int(N))
    return out


def xext_fft_from_tau(tau, N):
    out = np.zeros((16, N // 8, 18), dtype=np.uint64)
    lib().or_xext_fft_from_tau(_p(out), _p(fr(tau)), ctypes.c_int(N))
    return out


def fk20_shadow(poly, tau):
    poly = np.ascontiguousarray(poly, dtype=np.uint64)
    N = poly.shape[0]
    out = np.zeros((N // 8, 4), dtype=np.uint64)
    lib().or_fk20_shadow(_p(out), _p(poly), _p(fr(tau)), ctypes.c_int(N))
    return out


def kzg_quotient_at(poly, c, tau):
    poly = np.ascontiguousarray(poly, dtype=np.uint64)
    out = np.zeros(4, dtype=np.uint64)
    lib().or_kzg_quotient_at(_p(out), _p(poly), ctypes.c_int(poly.shape[0]),
                             _p(np.ascontiguousarray(c, dtype=np.uint64)), _p(fr(tau)))
    return out


def default_threads():
    return int(lib().or_default_threads())


def pool_threads():
    """Worker threads for the bulk checkers: the CPUs this process may run on, at most 16 (a
    GPU box's per-GPU CPU share, whatever the machine's total)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


# ---- bulk closed forms ----

def gen_mul_affine(scalars, nthreads=0):
    """Canonical affine [s]G (shape (..., 12), infinity (0, 0)) for fr scalars (..., 4), by
    fixed-base windows over G."""
    s = np.ascontiguousarray(scalars, dtype=np.uint64)
    n = s.size // 4
    out = np.zeros(s.shape[:-1] + (12,), dtype=np.uint64)
    lib().or_g1_gen_mul_affine_n(_p(out), _p(s), ctypes.c_int64(n), ctypes.c_int(nthreads or pool_threads()))
    return out


def gen_mul_projective(scalars, zs, nthreads=0):
    """[s]G as a non-normalised projective representative: the fixed-base sum's (X, Y, Z)
    scaled by z (infinity (0, z, 0)); zs are non-zero standard-form Fp elements (..., 6)."""
    s = np.ascontiguousarray(scalars, dtype=np.uint64)
    z = np.ascontiguousarray(zs, dtype=np.uint64)
    n = s.size // 4
    assert z.size // 6 == n
    out = np.zeros(s.shape[:-1] + (18,), dtype=np.uint64)
    lib().or_g1_gen_mul_proj_n(_p(out), _p(s), _p(z), ctypes.c_int64(n), ctypes.c_int(nthreads or pool_threads()))
    return out


def match_affine(proj, aff, nthreads=0):
    """(mismatches, first mismatching index or -1) of projective points (..., 18) against
    expected affine points (..., 12), compared by cross-multiplication."""
    p = np.ascontiguousarray(proj, dtype=np.uint64)
    a = np.ascontiguousarray(aff, dtype=np.uint64)
    n = p.size // 18
    assert a.size // 12 == n
    first = ctypes.c_int64()
    bad = lib().or_g1p_match_affine_n(_p(p), _p(a), ctypes.c_int64(n), ctypes.c_int(nthreads or pool_threads()),
                                      ctypes.byref(first))
    return int(bad), int(first.value)


def fk20_shadow_rows(poly, tau, nthreads=0):
    """Scalar shadow s_fft (rows, N/8, 4) of polynomials (rows, N, 4): h_fft[r][k] = [s[r][k]]G."""
    poly = np.ascontiguousarray(poly, dtype=np.uint64)
    rows, N = poly.shape[0], poly.shape[1]
    out = np.zeros((rows, N // 8, 4), dtype=np.uint64)
    lib().or_fk20_shadow_rows(_p(out), _p(poly), _p(fr(tau)), ctypes.c_int(N), ctypes.c_int(rows),
                              ctypes.c_int(nthreads or pool_threads()))
    return out


def msm_shadow(tc_fft, xs, nthreads=0):
    """Scalars of fk20_msm over bases [xs]G: out (rows, n2, 4) = sum_i tc_fft[:, i, :] * xs[i]."""
    tc = np.ascontiguousarray(tc_fft, dtype=np.uint64)
    xs = np.ascontiguousarray(xs, dtype=np.uint64)
    rows, n2 = tc.shape[0], tc.shape[2]
    assert tc.shape[1] == 16 and xs.shape[:2] == (16, n2)
    out = np.zeros((rows, n2, 4), dtype=np.uint64)
    lib().or_fk20_msm_shadow(_p(out), _p(tc), _p(xs), ctypes.c_int(n2), ctypes.c_int(rows),
                             ctypes.c_int(nthreads or pool_threads()))
    return out


def fk20_closed_form_affine(poly, tau, nthreads=0):
    """Expected h_fft of every row in canonical affine form (rows, N/8, 12) for a setup with
    secret tau: [s_k]G of the scalar shadow."""
    return gen_mul_affine(fk20_shadow_rows(poly, tau, nthreads), nthreads)


# ---- output formats ----

def reverse_bit_order(x):
    """x[..., bitrev(i), ...] along axis 0 of the point index: FK20Py list_to_reverse_bit_order
    (kzg_proofs.py:107-118) applied to the last point axis of x (shape (..., n, k))."""
    x = np.asarray(x)
    n = x.shape[-2]
    bits = n.bit_length() - 1
    idx = [int(format(i, "0%db" % bits)[::-1], 2) for i in range(n)]
    return x[..., idx, :]


def g1_compress(p):
    """48-byte compressed serialisation of g1p points (shape (..., 18)) -> uint8 (..., 48):
    big-endian canonical x; top byte flags 0x80 compressed, 0x40 infinity, 0x20 set when y is
    the larger of (y, p - y).  Infinity is 0xc0 followed by zeros."""
    p = np.ascontiguousarray(p, dtype=np.uint64)
    flat = p.reshape(-1, 18)
    aff = to_affine(flat)
    out = np.zeros((flat.shape[0], 48), dtype=np.uint8)
    for k in range(flat.shape[0]):
        if g1p_is_inf(flat[k]):
            out[k, 0] = 0xC0
            continue
        x, y = limbs_to_int(aff[k, :6]), limbs_to_int(aff[k, 6:])
        b = bytearray(x.to_bytes(48, "big"))
        b[0] |= 0x80 | (0x20 if y > P_MOD - y else 0)
        out[k] = np.frombuffer(bytes(b), dtype=np.uint8)
    return out.reshape(p.shape[:-1] + (48,))


def g1_decompress(b):
    """Inverse of g1_compress for one 48-byte string: affine (x, y) integers, or None for
    infinity.  y = (x^3 + 4)^((p+1)/4) (p = 3 mod 4), sign chosen by the 0x20 flag."""
    b = bytes(bytearray(b))
    assert b[0] & 0x80
    if b[0] & 0x40:
        return None
    x = int.from_bytes(bytes([b[0] & 0x1F]) + b[1:], "big")
    y = pow((x ** 3 + 4) % P_MOD, (P_MOD + 1) // 4, P_MOD)
    assert y * y % P_MOD == (x ** 3 + 4) % P_MOD
    if (y > P_MOD - y) != bool(b[0] & 0x20):
        y = P_MOD - y
    return x, y


# ---- affine G1 helpers for the pairing-based verifier (oracle/pairing.py) ----

def g1_affine(p):
    """g1p (18,) -> (x, y) integers, or None for infinity."""
    if g1p_is_inf(p):
        return None
    a = to_affine(p)
    return limbs_to_int(a[:6]), limbs_to_int(a[6:])


def g1p_from_affine(a):
    return g1p_inf() if a is None else g1p_from_ints(a[0], a[1], 1)


def g1_affine_add(a, b):
    return g1_affine(g1p_op("add", g1p_from_affine(a), g1p_from_affine(b)))


def g1_affine_mul(a, k):
    return g1_affine(g1p_op("mul", g1p_from_affine(a), fr(k % R_MOD)))


def fk20_proof_coset(k, N):
    """Coset shift of h_fft[k] for polynomial length N: omega_{2N}^k (the proof covers the 16
    points x * omega_16^i, the roots of X^16 - x^16)."""
    return pow(pow(5, (R_MOD - 1) // (2 * N), R_MOD), k, R_MOD)


def poly_eval(coeffs, x):
    acc = 0
    for c in reversed([limbs_to_int(v) for v in np.asarray(coeffs)]):
        acc = (acc * x + c) % R_MOD
    return acc
