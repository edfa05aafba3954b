"""KZG commitments of rows (fk20_commit_setup_load + fk20_commit_rows) against the oracle:
the closed form [f(tau)]G for an SRS with secret tau, and the direct linear combination of
FK20Py commit_to_poly (kzg_proofs.py:159) for non-normalised SRS points and a point at
infinity."""
import numpy as np
import pytest

import oracle as O
from fk20_amd import lib as L
from fk20_amd import synthetic as S
from fk20_amd.prover import Prover

pytestmark = pytest.mark.gpu


def aff(pts):
    return O.to_affine(np.ascontiguousarray(pts, dtype=np.uint64).reshape(-1, 18))


@pytest.fixture(scope="module")
def prover():
    p = Prover()
    yield p
    p.close()


@pytest.mark.parametrize("N,rows,window", [(1024, 70, 8), (2048, 65, 11), (4096, 64, 8)])
def test_commit_matches_closed_form(prover, N, rows, window):
    tau = S.random_secret(700 + N)
    srs = O.setup_srs(tau, N)
    poly = S.random_polys(rows, N, 701 + N)
    poly[1] = 0
    sid = prover.load_commit_setup(srs, N, window)
    got = prover.commit(sid, poly)
    prover.unload_setup(sid)
    assert O.g1p_is_inf(got[1])
    for r in [0, 2, rows - 1] + list(range(3, rows, 9)):
        assert np.array_equal(aff(got[r]), aff(O.commit_closed_form(poly[r], tau))), r


def test_commit_edge_setup_points(prover):
    """Non-normalised SRS points (random Z) and an SRS point at infinity: equal to the direct
    linear combination."""
    N, rows = 1024, 3
    srs = O.setup_srs(S.random_secret(711), N)
    z = S.random_fp_nonzero(N, 712)
    for i in range(0, N, 7):  # (X z, Y z, Z z) is the same point
        zi = O.limbs_to_int(z[i])
        for c in range(3):
            v = O.limbs_to_int(srs[i, 6 * c:6 * c + 6]) * zi % O.P_MOD
            srs[i, 6 * c:6 * c + 6] = [(v >> (64 * k)) & (2**64 - 1) for k in range(6)]
    srs[5] = O.g1p_inf()
    poly = S.random_polys(rows, N, 713)
    sid = prover.load_commit_setup(srs, N, 8)
    got = prover.commit(sid, poly)
    prover.unload_setup(sid)
    want = O.commit(poly, srs)
    assert np.array_equal(aff(got), aff(want))


def test_setup_kinds_are_not_interchangeable(prover):
    N = 1024
    tau = S.random_secret(721)
    cid = prover.load_commit_setup(O.setup_srs(tau, N), N, 8)
    fid = prover.load_setup(O.xext_fft_from_tau(tau, N), N, 8)
    buf = L.DeviceBuffer(N * 32)
    out = L.DeviceBuffer((N // 8) * 144)
    lib = L.lib()
    assert lib.fk20_prove_rows(prover.ctx, cid, out.ptr, buf.ptr, 1, None) == -1
    assert lib.fk20_commit_rows(prover.ctx, fid, out.ptr, buf.ptr, 1, None) == -1
    for b in (buf, out):
        b.free()
    prover.unload_setup(cid)
    prover.unload_setup(fid)
