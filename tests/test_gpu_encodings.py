"""Encodings the reference ABI allows besides the canonical one (FK20-CUDA fr.cuh:24 and
fp.cuh:14: standard form, values at or above the modulus allowed; g1.cuh:14-30: projective
points with any Z, and every (0, y != 0, 0) is the point at infinity).  Every entry point must
treat them as the values they encode: each test runs the same data canonically encoded and
re-encoded, and compares canonical affine points / residues."""
import numpy as np
import pytest

import oracle as O
from fk20_amd import ops
from fk20_amd import synthetic as S
from fk20_amd.prover import Prover

pytestmark = pytest.mark.gpu

R = O.R_MOD
P = O.P_MOD


def aff(pts):
    return O.to_affine(np.ascontiguousarray(pts, dtype=np.uint64).reshape(-1, 18))


def residues(a, mod):
    a = np.asarray(a)
    return [O.limbs_to_int(x) % mod for x in a.reshape(-1, a.shape[-1])]


def fr_plus_r(a):
    """Each Fr element v < r re-encoded as v + r (< 2^256 for every v < r)."""
    a = np.asarray(a, dtype=np.uint64)
    flat = [O.int_to_limbs(O.limbs_to_int(x) + R, 4) for x in a.reshape(-1, 4)]
    return np.array(flat, dtype=np.uint64).reshape(a.shape)


def fp_lift(v, rng):
    """v + k p for a random k in [1, 8] (v + 8p < 9p < 2^384)."""
    return v + int(rng.integers(1, 9)) * P


def g1_reencode(pts, seed, infinities=()):
    """Points with every coordinate lifted by a multiple of p; the indices in `infinities`
    become (0, y, 0) with a random y in [2, p) (lifted as well).  Returns (reencoded,
    canonical) where canonical holds the same points with (0, 1, 0) for those infinities."""
    rng = np.random.default_rng(seed)
    pts = np.array(pts, dtype=np.uint64).reshape(-1, 18)
    canon = pts.copy()
    out = np.zeros_like(pts)
    for i, p in enumerate(pts):
        x, y, z = (O.limbs_to_int(p[6 * k:6 * k + 6]) for k in range(3))
        if i in infinities:
            x, y, z = 0, 2 + int(rng.integers(0, 2 ** 62)) * 12345 % (P - 2), 0
            canon[i] = O.g1p_inf()
        out[i] = np.concatenate([O.fp(fp_lift(c, rng)) for c in (x, y, z)])
    return out, canon


def random_points(count, seed):
    G = O.g1p_gen()
    return np.array([O.g1p_op("mul", G, k) for k in S.random_fr(count, seed)])


def test_fr_transforms_and_toeplitz():
    x = S.random_fr(4 * 512, 3).reshape(4, 512, 4)
    xn = fr_plus_r(x)
    for inv in (False, True):
        assert residues(ops.fr_fft(xn, inverse=inv), R) == residues(ops.fr_fft(x, inverse=inv), R)
    poly = S.random_polys(2, 1024, 4)
    polyn = fr_plus_r(poly)
    assert residues(ops.poly2toeplitz_coefficients(polyn), R) == residues(ops.poly2toeplitz_coefficients(poly), R)
    assert residues(ops.poly2toeplitz_coefficients_fft(polyn), R) == \
        residues(ops.poly2toeplitz_coefficients_fft(poly), R)


def test_field_ops():
    a, b = S.random_fr(64, 5), S.random_fr(64, 6)
    for op in ("add", "sub", "mul"):
        assert residues(ops.fr_op(op, fr_plus_r(a), fr_plus_r(b)), R) == residues(ops.fr_op(op, a, b), R)
    rng = np.random.default_rng(7)
    xs = [int(rng.integers(0, 2 ** 63)) ** 6 % P for _ in range(64)]
    ys = [int(rng.integers(0, 2 ** 63)) ** 6 % P for _ in range(64)]
    fa, fb = np.array([O.fp(v) for v in xs]), np.array([O.fp(v) for v in ys])
    fan = np.array([O.fp(fp_lift(v, rng)) for v in xs])
    fbn = np.array([O.fp(fp_lift(v, rng)) for v in ys])
    for op in ("add", "sub", "mul"):
        assert residues(ops.fp_op(op, fan, fbn), P) == residues(ops.fp_op(op, fa, fb), P)


def test_g1_helpers():
    pts = random_points(32, 8)
    ren, canon = g1_reencode(pts, 9, infinities=(0, 5))
    assert np.array_equal(ops.to_affine(ren), O.to_affine(canon))
    assert np.array_equal(ops.compress(ren), ops.compress(canon))
    eq = ops.g1p_op("eq", ren, canon)[:, 0]
    assert eq.tolist() == [1] * 32
    assert ops.g1p_op("is_point", ren)[:, 0].tolist() == [1] * 32
    k = S.random_fr(32, 10)
    assert np.array_equal(aff(ops.g1p_op("mul", ren, k=fr_plus_r(k))), aff(ops.g1p_op("mul", canon, k=k)))
    assert np.array_equal(aff(ops.g1p_op("add", ren, ren[::-1].copy())),
                          aff(ops.g1p_op("add", canon, canon[::-1].copy())))


@pytest.mark.parametrize("n", [128, 512])
def test_g1_transforms(n):
    pts = random_points(n, 11 + n)
    ren, canon = g1_reencode(pts, 12, infinities=(1, n // 2, n - 1))
    for inv in (False, True):
        got = ops.g1p_fft(ren[None], inverse=inv)
        assert np.array_equal(aff(got), aff(ops.g1p_fft(canon[None], inverse=inv))), (n, inv)
    assert np.array_equal(aff(ops.hext_fft2h_fft(ren[None])), aff(ops.hext_fft2h_fft(canon[None])))


@pytest.fixture(scope="module")
def case():
    N = 1024
    tau = S.random_secret(13)
    xext = O.xext_fft_from_tau(tau, N)
    xn, _ = g1_reencode(xext, 14)
    poly = S.random_polys(3, N, 15)
    return N, xext, xn.reshape(xext.shape), poly


def test_msm(case):
    N, xext, xn, poly = case
    tc_fft = ops.poly2toeplitz_coefficients_fft(poly)
    exp = ops.msm(tc_fft, xext)
    # scalars lifted by r (bit 255 set where v + r >= 2^255) and bases re-encoded
    got = ops.msm(fr_plus_r(tc_fft), xn)
    assert np.array_equal(aff(got), aff(exp))
    assert np.array_equal(aff(exp[0]), aff(O.msm(tc_fft[0], xext)))


def test_poly2h_fft_plain_and_batched(case):
    N, xext, xn, poly = case
    polyn = fr_plus_r(poly)
    exp = O.poly2h_fft(poly[:1], xext)
    plain = ops.poly2h_fft(polyn, xn)
    assert np.array_equal(aff(plain[0]), aff(exp[0]))
    with Prover() as p:
        sid = p.load_setup(xn, N, 8)
        got = p.prove(sid, polyn)
    assert np.array_equal(aff(got), aff(plain))


def test_setup2xext_fft():
    N = 1024
    tau = S.random_secret(16)
    srs = O.setup_srs(tau, N + 1)
    srsn, _ = g1_reencode(srs, 17)
    assert np.array_equal(aff(ops.setup2xext_fft(srsn, N)), aff(O.xext_fft_from_tau(tau, N)))


@pytest.mark.parametrize("window", [8, 9, 13])
def test_commit_non_canonical_coefficients(window):
    """fk20_commit_rows with coefficients re-encoded as v + r and the edge values r, 2^255,
    2^256 - 1 (their residues mod r are what the commitment must use).  At window 8 the signed
    windows end exactly at bit 256, so an unreduced value >= 2^255 would carry out of them."""
    N, rows = 1024, 3
    tau = S.random_secret(7400 + window)
    srs = O.setup_srs(tau, N)
    poly = S.random_polys(rows, N, 7401)
    enc = fr_plus_r(poly)
    edge = [R, 1 << 255, (1 << 256) - 1, R - 1, (1 << 256) - R]
    for k, v in enumerate(edge):
        enc[2, k] = O.int_to_limbs(v, 4)
        poly[2, k] = O.int_to_limbs(v % R, 4)
    with Prover() as p:
        sid = p.load_commit_setup(srs, N, window)
        got = p.commit(sid, enc)
        want = p.commit(sid, poly)
    assert np.array_equal(aff(got), aff(want))
    for r in range(rows):
        assert np.array_equal(aff(got[r]), aff(O.commit_closed_form(poly[r], tau))), r
