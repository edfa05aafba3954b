"""CPU check of the device arithmetic headers (fp.hip.h, fr.hip.h, g1.hip.h) compiled by g++
from the same source the HIP kernels inline.  Exercises the lazy-reduction bounds with
random and extreme inputs against Python integers and the oracle.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from fk20_amd import synthetic as S

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native", "arith_host.cpp")
OUT = os.path.join(HERE, "native", "_build", "libarith_host.so")
P, R = O.P_MOD, O.R_MOD


@pytest.fixture(scope="module")
def L():
    deps = [SRC] + [os.path.join(HERE, "..", "fk20-cuda_amd", "csrc", f) for f in
                    ("fp.hip.h", "fr.hip.h", "g1.hip.h", "constants.hip.h")]
    if not os.path.exists(OUT) or os.path.getmtime(OUT) < max(os.path.getmtime(d) for d in deps):
        os.makedirs(os.path.dirname(OUT), exist_ok=True)
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", OUT, SRC], check=True)
    lib = ctypes.CDLL(OUT)
    return lib


def p64(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def fp_call(L, op, a, b, reps=1):
    z = np.zeros(6, dtype=np.uint64)
    av, bv = O.fp(a), (O.fp(b) if b is not None else None)
    L.host_fp_op(op, p64(z), p64(av), p64(bv) if bv is not None else None, reps)
    return O.limbs_to_int(z)


def test_fp_ops_random_and_edges(L):
    rng = np.random.default_rng(11)
    vals = [0, 1, 2, P - 1, P, P + 1, 2**384 - 1, 2**381, 9 * P] + [int.from_bytes(rng.bytes(48), "little") for _ in range(300)]
    for i, a in enumerate(vals):
        b = vals[(i * 7 + 3) % len(vals)]
        assert fp_call(L, 2, a, b) == a * b % P
        assert fp_call(L, 0, a, b) == (a + b) % P
        assert fp_call(L, 1, a, b) == (a - b) % P
        assert fp_call(L, 3, a, None) == a * a % P
        assert fp_call(L, 4, a, None) == (-a) % P
        assert fp_call(L, 6, a, None) == a % P
    for a in vals[:40]:
        assert fp_call(L, 5, a, None) == (pow(a, P - 2, P))


def test_fp_lazy_chains(L):
    """Repeated add without reduction stays exact (bounds grow, still < 2^12 p)."""
    rng = np.random.default_rng(12)
    for _ in range(50):
        a = int.from_bytes(rng.bytes(48), "little") % P
        b = int.from_bytes(rng.bytes(48), "little") % P
        assert fp_call(L, 0, a, b, reps=6) == (a + 6 * b) % P
        assert fp_call(L, 2, a, b, reps=20) == a * pow(b, 20, P) % P
        assert fp_call(L, 3, a, None, reps=20) == pow(a, 2**20, P)


def _limbs_int(v):
    return sum(int(x) << (29 * k) for k, x in enumerate(v))


def test_fp_mul_limb_extremes(L):
    """fe_mul and fe_mul2 (two products, one reduction) are exact with limbs at their maxima:
    lower limbs 2^29 - 1, top limbs at the 2^12 p value bound, and one operand carrying
    unpropagated sums (fe_add_nn: lower limbs up to 2^30 - 1), where the column sums peak."""
    rng = np.random.default_rng(13)
    RINV = pow(2, -406, P)
    top = (2**12 * P) >> (29 * 13)
    full = [(1 << 29) - 1] * 13 + [top]
    nn = [(1 << 30) - 1] * 13 + [2 * top]

    def rand(bits):
        return [int(x) for x in rng.integers(0, 1 << bits, 13)] + [int(rng.integers(0, top + 1))]
    quads = [[nn, full, full, full], [full, full, full, full], [full, full, nn, full]]
    quads += [[rand(30), rand(29), rand(29), rand(29)] for _ in range(60)]
    quads += [[rand(29), rand(29), rand(29), rand(30)] for _ in range(60)]
    z = np.zeros(14, dtype=np.uint32)
    zp = z.ctypes.data_as(ctypes.c_void_p)
    for q in quads:
        ops = [np.array(c, dtype=np.uint32) for c in q]
        a, b, c, d = [_limbs_int(o) for o in ops]
        ptr = [o.ctypes.data_as(ctypes.c_void_p) for o in ops]
        L.host_fp_mul2_raw(zp, *ptr)
        zi = _limbs_int(z)
        assert zi % P == (a * b + c * d) * RINV % P
        assert zi < 2 * P and all(int(x) < (1 << 29) for x in z[:13])
        L.host_fp_mul_raw(zp, ptr[0], ptr[1])
        zi = _limbs_int(z)
        assert zi % P == a * b * RINV % P
        assert zi < 2 * P and all(int(x) < (1 << 29) for x in z[:13])


def test_fr_mul(L):
    xs = S.random_fr(200, 5)
    ys = S.random_fr(200, 6)
    for x, y in zip(xs, ys):
        z = np.zeros(4, dtype=np.uint64)
        L.host_fr_mul(p64(z), p64(x), p64(y))
        assert O.limbs_to_int(z) == O.limbs_to_int(x) * O.limbs_to_int(y) % R


def test_g1_formulas_against_oracle(L):
    G = O.g1p_gen()
    pts = [O.g1p_op("mul", G, O.fr(k)) for k in (1, 2, 3, 12345, R - 1)] + [O.g1p_inf()]
    for p in pts:
        for q in pts:
            r = np.zeros(18, dtype=np.uint64)
            L.host_g1_add(p64(r), p64(p), p64(q))
            assert O.g1p_eq(r, O.g1p_op("add", p, q))
        r = np.zeros(18, dtype=np.uint64)
        L.host_g1_dbl(p64(r), p64(p))
        assert O.g1p_eq(r, O.g1p_op("dbl", p))
    # mixed addition: affine right operand (never infinity), including q = +-p
    for q in pts[:-1]:
        qa = O.to_affine(q)
        for p in pts:
            r = np.zeros(18, dtype=np.uint64)
            L.host_g1_add_aff(p64(r), p64(p), p64(qa))
            assert O.g1p_eq(r, O.g1p_op("add", p, q))
    for k in S.random_fr(4, 9):
        r = np.zeros(18, dtype=np.uint64)
        L.host_g1_mul(p64(r), p64(G), p64(k))
        assert O.g1p_eq(r, O.g1p_op("mul", G, k))


def test_g1_chain_lazy(L):
    """dbl/add/neg chains: b_{i+1} = (2 a_i + b_i) - b_i, a_{i+1} = 2 a_i + b_i."""
    G = O.g1p_gen()
    r = np.zeros(18, dtype=np.uint64)
    L.host_g1_chain(p64(r), p64(G), 30)
    # a_0 = b_0 = G; a_{i+1} = 2 a_i + b_i, b_{i+1} = a_{i+1} - b_i
    a, b = 1, 1
    for _ in range(30):
        a = 2 * a + b
        b = a - b
    assert O.g1p_eq(r, O.g1p_op("mul", G, O.fr(b % R)))


def test_g1_compress_against_oracle(L):
    """The device serialisation (g1_compress) equals the oracle's for G, -G, infinity and
    random multiples, from non-normalised projective inputs."""
    G = O.g1p_gen()
    pts = [G, O.g1p_op("neg", G), O.g1p_inf()] + [O.g1p_op("mul", G, k) for k in S.random_fr(20, 31)]
    for p in pts:
        p = np.ascontiguousarray(p, dtype=np.uint64)
        out = np.zeros(48, dtype=np.uint8)
        L.host_g1_compress(out.ctypes.data_as(ctypes.c_void_p), p64(p))
        assert bytes(out) == bytes(O.g1_compress(p))
