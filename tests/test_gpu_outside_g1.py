"""Inputs outside the order-r subgroup G1.

On a point P outside G1, [a]P and [b]P differ whenever a != b as integers, even for a = b
(mod r), so a transform of such points is fixed by its exact sequence of multipliers.  The
library checks the transforms' inputs and recomputes a row holding such a point in the
reference's order of operations (g1p_fft.cu:25-330, fk20_hext_fft2h_fft.cu:27-70); the
oracle follows the same order, forward and inverse, and is pinned on G1 by the reference's
KATs.  Points outside G1 are random points of E(Fp) (y^2 = x^3 + 4); no reference vector
covers them.  G1 results are compared as canonical affine points.
"""
import numpy as np
import pytest

import oracle as O
from fk20_amd import lib as L
from fk20_amd import ops
from fk20_amd import synthetic as S
from fk20_amd.prover import Prover

pytestmark = pytest.mark.gpu


def aff(pts):
    return O.to_affine(np.ascontiguousarray(pts, dtype=np.uint64).reshape(-1, 18))


def g1_points(count, seed):
    G = np.tile(O.g1p_gen(), (count, 1))
    return ops.g1p_op("mul", G, k=S.random_fr(count, seed))


def mixed_rows(n, seed):
    """Three rows of n points: in G1; every point outside G1; in G1 but for position 5."""
    x = g1_points(3 * n, seed).reshape(3, n, 18)
    x[1] = S.curve_points(n, seed + 1)
    x[2, 5] = S.curve_points(1, seed + 2)[0]
    return x


@pytest.mark.parametrize("n", [128, 256, 512])
def test_g1_transforms_outside_g1(n):
    x = mixed_rows(n, 500 + n)
    fwd = ops.g1p_fft(x)
    inv = ops.g1p_fft(x, inverse=True)
    for r in range(3):
        assert np.array_equal(aff(fwd[r]), aff(O.g1p_fft(x[r]))), r
        assert np.array_equal(aff(inv[r]), aff(O.g1p_fft(x[r], inverse=True))), r
    # in place (the output aliases the input, as g1p_fft.cu:17 allows)
    assert np.array_equal(aff(ops.g1p_fft(x, inplace=True)), aff(fwd))
    assert np.array_equal(aff(ops.g1p_fft(x, inverse=True, inplace=True)), aff(inv))


def test_order_of_operations_matters_outside_g1():
    """The check above can tell orders apart: outside G1 the inverse transform differs from
    the forward transform with inverse roots times n^-1 (the same map on G1)."""
    n = 128
    x = S.curve_points(n, 520)
    inv = aff(ops.g1p_fft(x[None], inverse=True)[0])
    # forward DFT with w^-1: out[k] = fft(x)[-k mod n], then [n^-1]
    fwd = ops.g1p_fft(x[None])[0]
    ninv = O.fr(pow(n, -1, O.R_MOD))
    alt = np.array([O.g1p_op("mul", fwd[(-k) % n], ninv) for k in range(n)])
    assert not np.array_equal(inv, aff(alt))
    g = g1_points(n, 521)
    inv_g = aff(ops.g1p_fft(g[None], inverse=True)[0])
    fwd_g = ops.g1p_fft(g[None])[0]
    alt_g = np.array([O.g1p_op("mul", fwd_g[(-k) % n], ninv) for k in range(n)])
    assert np.array_equal(inv_g, aff(alt_g))


def test_hext_fft2h_fft_outside_g1():
    n2 = 128
    x = mixed_rows(n2, 530)
    got = ops.hext_fft2h_fft(x)
    for r in range(3):
        assert np.array_equal(aff(got[r]), aff(O.hext_fft2h_fft(x[r]))), r


def test_poly2h_fft_setup_outside_g1():
    """A setup with one base outside G1: every row is recomputed from the MSM output on in the
    reference's order (the MSM's scalars are canonical, as the oracle's)."""
    N, rows = 1024, 2
    xext = O.xext_fft_from_tau(S.random_secret(540), N).copy()
    xext[3, 7] = S.curve_points(1, 541)[0]
    poly = S.random_polys(rows, N, 542)
    got = ops.poly2h_fft(poly, xext)
    exp = O.poly2h_fft(poly, xext)
    assert np.array_equal(aff(got), aff(exp))


def test_setup_outside_g1_rejected():
    """The batched prover folds 1/n2 into its window tables and runs endomorphism-split
    twiddles: a setup outside G1 is refused, and a good one still loads afterwards."""
    N = 1024
    xext = O.xext_fft_from_tau(S.random_secret(550), N).copy()
    good = xext.copy()
    xext[0, 1] = S.curve_points(1, 551)[0]
    with Prover() as p:
        with pytest.raises(L.Fk20Error) as e:
            p.load_setup(xext, N, 8)
        assert e.value.status == -1
        sid = p.load_setup(good, N, 8)
        assert sid >= 0


def test_commit_setup_outside_g1():
    """Commitments multiply each SRS point by its coefficient only, so any point of E(Fp)
    will do; one of small order (here 3) is refused, its multiples reaching infinity inside
    the affine window tables."""
    N = 1024
    srs = O.setup_srs(S.random_secret(560), N).copy()
    srs[9] = S.curve_points(1, 561)[0]
    poly = S.random_polys(2, N, 562)
    with Prover() as p:
        sid = p.load_commit_setup(srs, N, 8)
        got = p.commit(sid, poly)
    assert np.array_equal(aff(got), aff(O.commit(poly, srs)))
    # a point of order 3: [r h / 3] Q
    q = S.curve_points(1, 563)[0]
    t = O.g1p_op("mul", q, O.fr(O.R_MOD))
    h3 = 0x396c8c005555e1568c00aaab0000aaab // 3
    t = O.g1p_op("add", O.g1p_op("mul", O.g1p_op("mul", t, O.fr(h3 >> 128)), O.fr(1 << 128)),
                 O.g1p_op("mul", t, O.fr(h3 & ((1 << 128) - 1))))
    assert not O.g1p_is_inf(t) and O.g1p_is_inf(O.g1p_op("mul", t, O.fr(3)))
    srs[9] = t
    with Prover() as p:
        with pytest.raises(L.Fk20Error) as e:
            p.load_commit_setup(srs, N, 8)
        assert e.value.status == -1
