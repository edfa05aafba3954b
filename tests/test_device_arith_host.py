"""CPU check of the device arithmetic headers (fp.hip.h, fr.hip.h, g1.hip.h) compiled by g++
from the same source the HIP kernels inline.  Exercises the lazy-reduction bounds with
random and extreme inputs against Python integers and the oracle.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from fk20_amd import synthetic as S

P, R = O.P_MOD, O.R_MOD


@pytest.fixture(scope="module")
def L(arith_host):
    return arith_host


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


# ---------------------------------------------------------------- Jacobian ladder

def _affine_ints(p):
    a = O.to_affine(np.ascontiguousarray(p, dtype=np.uint64))
    return O.limbs_to_int(a[:6]), O.limbs_to_int(a[6:])


def _jac(x, y, z):
    """Jacobian slots (X, Y, Z) = (x z^2, y z^3, z) of the affine point (x, y), standard form."""
    return np.concatenate([O.fp(x * z * z % P), O.fp(y * z * z * z % P), O.fp(z)])


def _jac_affine(r):
    X, Y, Z = (O.limbs_to_int(r[6 * i:6 * i + 6]) for i in range(3))
    zi = pow(Z, P - 2, P)
    return X * zi * zi % P, Y * zi * zi * zi % P


def test_jacobian_dbl_and_madd(L):
    """g1j_dbl (repeated, lazy bounds carried between steps) and g1j_madd against the oracle's
    scalar multiples, from Jacobian representatives with random Z."""
    G = O.g1p_gen()
    rng = np.random.default_rng(41)
    for a, b in [(1, 2), (5, 3), (R - 7, 11), (12345, R - 12344 - 5)] + [(int(x), int(y)) for x, y in
                                                                       rng.integers(1, 2**62, (10, 2))]:
        za = int.from_bytes(rng.bytes(48), "little") % P or 1
        pa = _affine_ints(O.g1p_op("mul", G, O.fr(a)))
        qb = _affine_ints(O.g1p_op("mul", G, O.fr(b)))
        r = np.zeros(18, dtype=np.uint64)
        reps = int(rng.integers(1, 40))
        L.host_g1j_dbl(p64(r), p64(_jac(*pa, za)), reps)
        assert _jac_affine(r) == _affine_ints(O.g1p_op("mul", G, O.fr(a * 2**reps % R)))
        L.host_g1j_madd(p64(r), p64(_jac(*pa, za)), p64(np.concatenate([O.fp(qb[0]), O.fp(qb[1])])))
        assert _jac_affine(r) == _affine_ints(O.g1p_op("mul", G, O.fr((a + b) % R)))


def _programs():
    from fk20_amd import prover
    for which, count in ((0, 512), (1, 512), (2, 10)):
        for idx in range(1 if which < 2 else 0, count):
            yield which, idx, prover.twiddle_program(which, idx)


def _program_scalar(which, idx):
    w = pow(5, (R - 1) // 512, R)
    if which == 2:
        return pow(2, -idx, R)
    return pow(w, idx if which == 0 else -idx, R)


def test_ladder_runs_every_program(L):
    """The transform kernels' ladder (co-Z table, Jacobian doublings, mixed additions; the
    source the kernels inline) computes k P for every program the library builds -- w^+-i,
    i = 1..511, and 2^-j -- on random points given as non-normalised projective inputs."""
    rng = np.random.default_rng(42)
    progs = list(_programs())
    s = S.random_fr(len(progs), 43)
    zs = np.array([O.fp(int.from_bytes(rng.bytes(48), "little") % (P - 1) + 1) for _ in progs])
    pts = O.gen_mul_projective(s, zs)
    ks = [O.fr(_program_scalar(w, i) * O.limbs_to_int(si) % R) for (w, i, _), si in zip(progs, s)]
    exp = O.gen_mul_affine(np.array(ks))
    out = np.zeros((len(progs), 18), dtype=np.uint64)
    for n, (w, i, ops) in enumerate(progs):
        L.host_tw_mul(p64(out[n]), p64(pts[n]), ops.ctypes.data_as(ctypes.c_void_p))
    bad, first = O.match_affine(out, exp)
    assert bad == 0, "program %s differs" % (progs[first][:2],)


def test_ladder_infinity_and_structured_inputs(L):
    """Infinity in any (0 : y : 0) form comes back as infinity; G, -G and small multiples (the
    table entries' own scalars) give exact results."""
    from fk20_amd import prover
    G = O.g1p_gen()
    for which, idx in ((0, 1), (1, 128), (0, 384), (2, 9)):
        ops = prover.twiddle_program(which, idx)
        k = _program_scalar(which, idx)
        for y in (1, 7, P - 1):
            inf = np.concatenate([O.fp(0), O.fp(y), O.fp(0)])
            r = np.zeros(18, dtype=np.uint64)
            L.host_tw_mul(p64(r), p64(inf), ops.ctypes.data_as(ctypes.c_void_p))
            assert O.g1p_is_inf(r)
        for m in (1, 2, 3, 15, R - 1, R - 15):
            p = O.g1p_op("mul", G, O.fr(m))
            r = np.zeros(18, dtype=np.uint64)
            L.host_tw_mul(p64(r), p64(p), ops.ctypes.data_as(ctypes.c_void_p))
            assert O.g1p_eq(r, O.g1p_op("mul", G, O.fr(k * m % R)))


@pytest.mark.parametrize("w", [3, 4])
def test_ladder_other_widths(L, w):
    """The schedule variants' narrower tables: programs built at wNAF width 3 and 4 (2 and 4
    odd multiples) run by the ladder of that width, a sample of twiddles and every 2^-j."""
    from fk20_amd import prover
    rng = np.random.default_rng(44 + w)
    progs = [(which, idx) for which in (0, 1) for idx in (1, 2, 3, 64, 128, 129, 255, 256, 384, 511)]
    progs += [(2, j) for j in range(10)]
    s = S.random_fr(len(progs), 45 + w)
    zs = np.array([O.fp(int.from_bytes(rng.bytes(48), "little") % (P - 1) + 1) for _ in progs])
    pts = O.gen_mul_projective(s, zs)
    exp = O.gen_mul_affine(np.array([O.fr(_program_scalar(a, b) * O.limbs_to_int(si) % R)
                                     for (a, b), si in zip(progs, s)]))
    out = np.zeros((len(progs), 18), dtype=np.uint64)
    for n, (a, b) in enumerate(progs):
        ops = prover.twiddle_program(a, b, w)
        L.host_tw_mul_w(p64(out[n]), p64(pts[n]), ops.ctypes.data_as(ctypes.c_void_p), w)
    bad, first = O.match_affine(out, exp)
    assert bad == 0, "program %s differs" % (progs[first],)


# cofactor of G1 in E(Fp): h = 3 * 11^2 * 10177^2 * 859267^2 * 52437899^2
H_COF = 0x396c8c005555e1568c00aaab0000aaab


def _mul_big(p, k):
    """[k]p for an integer k of any size (the oracle's double-and-add takes 256 bits)."""
    acc = O.g1p_inf()
    chunks = []
    while k:
        chunks.append(k & ((1 << 128) - 1))
        k >>= 128
    for c in reversed(chunks):
        acc = O.g1p_op("mul", acc, O.fr(1 << 128))
        acc = O.g1p_op("add", acc, O.g1p_op("mul", p, O.fr(c)))
    return acc


def test_subgroup_membership(L):
    """The kernels' G1 membership test (phi^2(P) + [z^2]P = 0) against the definition [r]P = 0:
    G1 points, infinity, random points of E(Fp) (outside G1), their cofactor-cleared
    multiples [h]P (in G1), points of small order (3, 11, 121) and G1 + small-order sums."""
    G = O.g1p_gen()
    pts = [G, O.g1p_inf(), O.g1p_op("mul", G, O.fr(R - 1))]
    pts += [O.g1p_op("mul", G, s) for s in S.random_fr(3, 61)]
    curve = list(S.curve_points(6, 62))
    pts += curve
    pts += [_mul_big(q, H_COF) for q in curve[:3]]
    small = []
    for q in curve[3:]:
        rq = O.g1p_op("mul", q, O.fr(R))
        for m in (3, 11, 121):
            t = _mul_big(rq, H_COF // m)
            if not O.g1p_is_inf(t):
                small.append(t)
    assert len(small) >= 3
    pts += small
    pts += [O.g1p_op("add", G, t) for t in small[:2]]
    for p in pts:
        want = O.g1p_is_inf(O.g1p_op("mul", p, O.fr(R)))
        assert L.host_g1_in_subgroup(p64(np.ascontiguousarray(p))) == int(want)
    # the random curve points really are outside G1 and their [h] multiples inside
    assert not any(O.g1p_is_inf(O.g1p_op("mul", q, O.fr(R))) for q in curve)


def test_w128_short_form(L):
    """[w^+-128]P by 63 doublings of P and phi^2 (w^128 = |z|^3 = -|z| lambda^2 on G1) equals
    the oracle's multiple, on random non-normalised points, on G, -G, small multiples and on
    infinity in (0 : y : 0) form."""
    rng = np.random.default_rng(71)
    w = pow(5, (R - 1) // 512, R)
    s = S.random_fr(24, 72)
    zs = np.array([O.fp(int.from_bytes(rng.bytes(48), "little") % (P - 1) + 1) for _ in s])
    pts = list(O.gen_mul_projective(s, zs))
    G = O.g1p_gen()
    pts += [G, O.g1p_op("mul", G, O.fr(R - 1)), O.g1p_op("mul", G, O.fr(3)), O.g1p_op("mul", G, O.fr(2))]
    for w128, e in ((1, 128), (0, 384)):
        k = pow(w, e, R)
        for p in pts:
            r = np.zeros(18, dtype=np.uint64)
            L.host_tw_mul_z3(p64(r), p64(np.ascontiguousarray(p)), w128)
            assert O.g1p_eq(r, O.g1p_op("mul", p, O.fr(k)))
        for y in (1, 7, P - 1):
            inf = np.concatenate([O.fp(0), O.fp(y), O.fp(0)])
            r = np.zeros(18, dtype=np.uint64)
            L.host_tw_mul_z3(p64(r), p64(inf), w128)
            assert O.g1p_is_inf(r)
