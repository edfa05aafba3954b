"""The benchmark configuration itself (512 rows, N = 4096, one setup): the batched path
against the plain per-row sequence on every row, linearity of the proofs in the polynomial,
and closed-form spot checks."""
import numpy as np
import pytest

import oracle as O
from fk20_amd import ops
from fk20_amd import synthetic as S
from fk20_amd.prover import Prover

pytestmark = pytest.mark.gpu

R = O.R_MOD


def aff(pts):
    return O.to_affine(np.ascontiguousarray(pts, dtype=np.uint64).reshape(-1, 18))


@pytest.fixture(scope="module")
def full():
    N, rows = 4096, 512
    tau = S.random_secret(4096)
    xext = O.xext_fft_from_tau(tau, N)
    poly = S.random_polys(rows, N, 4097)
    p = Prover()
    sid = p.load_setup(xext, N, 8)
    yield N, rows, tau, xext, poly, p, sid
    p.close()


def test_batched_equals_plain_all_rows(full):
    N, rows, tau, xext, poly, p, sid = full
    got = p.prove(sid, poly)
    plain = ops.poly2h_fft(poly, xext)
    assert np.array_equal(aff(got), aff(plain))


def test_linearity(full):
    """h_fft(a + b) = h_fft(a) + h_fft(b) and h_fft(c a) = c h_fft(a), point by point."""
    N, rows, tau, xext, poly, p, sid = full
    a, b = poly[:16], poly[16:32]
    ab = poly[32:48].copy()
    ca = poly[48:64].copy()
    c = 0x1234567890ABCDEF1234567890ABCDEF % R
    for r in range(16):
        for j in range(N):
            x = O.limbs_to_int(a[r, j])
            y = O.limbs_to_int(b[r, j])
            ab[r, j] = O.fr((x + y) % R)
            ca[r, j] = O.fr(x * c % R)
    ha, hb, hab, hca = (p.prove(sid, q) for q in (a, b, ab, ca))
    summed = ops.g1p_op("add", ha.reshape(-1, 18), hb.reshape(-1, 18))
    assert np.array_equal(aff(summed), aff(hab))
    scaled = ops.g1p_op("mul", ha.reshape(-1, 18), k=np.tile(O.fr(c), (ha.size // 18, 1)))
    assert np.array_equal(aff(scaled), aff(hca))


def test_closed_form_rows(full):
    N, rows, tau, xext, poly, p, sid = full
    got = p.prove(sid, poly)
    G = O.g1p_gen()
    for r in (0, 255, 511):
        exp = np.array([O.g1p_op("mul", G, s) for s in O.fk20_shadow(poly[r], tau)])
        assert np.array_equal(aff(got[r]), aff(exp)), r
