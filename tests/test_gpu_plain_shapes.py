"""Drop-in entry points at ragged shapes: counts and row numbers that leave partial waves and
partial workgroups (65 = 64 + 1, 300 = 256 + 44), in place and out of place, against the oracle
and against each other (the batched path and the plain one)."""
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


def points(count, seed):
    """count points k G (k uniform), computed on the GPU: projective, Z not 1."""
    G = np.tile(O.g1p_gen(), (count, 1))
    return ops.g1p_op("mul", G, k=S.random_fr(count, seed))


@pytest.mark.parametrize("count,n", [(65, 128), (3, 512)])
def test_g1_transforms_ragged_and_in_place(count, n):
    x = points(count * n, 100 + count).reshape(count, n, 18)
    fwd = ops.g1p_fft(x)
    inv = ops.g1p_fft(x, inverse=True)
    # in place equals out of place
    assert np.array_equal(ops.g1p_fft(x, inplace=True), fwd)
    assert np.array_equal(ops.g1p_fft(x, inverse=True, inplace=True), inv)
    # round trips, and the first and last transforms against the oracle
    assert np.array_equal(aff(ops.g1p_fft(fwd, inverse=True)), aff(x))
    for r in (0, count - 1):
        assert np.array_equal(aff(fwd[r]), aff(O.g1p_fft(x[r])))
        assert np.array_equal(aff(inv[r]), aff(O.g1p_fft(x[r], inverse=True)))


@pytest.mark.parametrize("rows", [1, 5, 16, 17, 65, 300])
def test_msm_ragged_rows(rows):
    """1..16 rows take the term-parallel kernel (lanes = term x row), more rows the
    row-parallel one (lanes = rows, partial workgroups at 65 and 300)."""
    n2 = 128
    tc = S.random_fr(rows * 16 * n2, 200 + rows).reshape(rows, 16, n2, 4)
    bases = points(16 * n2, 201).reshape(16, n2, 18)
    got = ops.msm(tc, bases)
    for r in sorted({0, min(3, rows - 1), min(4, rows - 1), min(63, rows - 1), min(64, rows - 1), rows - 1}):
        assert np.array_equal(aff(got[r]), aff(O.msm(tc[r], bases))), r


def test_hext_fft2h_fft_and_poly2h_fft_ragged():
    N, rows = 1024, 65
    tau = S.random_secret(300)
    xext = S.device_setup(N, tau)
    xh = xext.download((16, N // 8, 18))
    poly = S.random_polys(rows, N, 301)
    plain = ops.poly2h_fft(poly, xh)
    with Prover() as p:
        sid = p.load_setup(xext, N, 8)
        assert np.array_equal(aff(p.prove(sid, poly)), aff(plain))
    xext.free()
    # the fused transform on its own, in place through the C ABI
    hext = points(rows * 128, 302).reshape(rows, 128, 18)
    out = ops.hext_fft2h_fft(hext)
    buf = L.DeviceBuffer.from_array(hext)
    L.call("fk20_hext_fft2h_fft", buf.ptr, buf.ptr, rows, 128, None)
    L.synchronize()
    assert np.array_equal(buf.download(hext.shape), out)
    buf.free()
    for r in (0, rows - 1):
        assert np.array_equal(aff(out[r]), aff(O.hext_fft2h_fft(hext[r])))
