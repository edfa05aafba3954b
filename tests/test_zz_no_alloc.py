"""fk20_poly2h_fft_noalloc (the allocation-free form of fk20_poly2h_fft, workspace in a fixed
device array) against fk20_poly2h_fft on the same inputs, bit-exact in the reference ABI's
standard-form projective output: two calls in a row on the shared array, the second with a
smaller input (fewer rows and a smaller N), so stale contents of the first call cannot leak
into the second; then the largest request that fits at N = 4096 and one row more, which is
rejected with FK20_ERANGE before any device work."""
import ctypes

import numpy as np
import pytest

import oracle as O
from fk20_amd import synthetic as S
from fk20_amd.lib import DeviceBuffer, call, lib, synchronize

pytestmark = pytest.mark.gpu

FK20_ERANGE = -6
SCRATCH = 64 << 20


def max_rows(N):
    n2 = N // 8
    return (SCRATCH - 2304 * n2) // (1344 * n2)


def run(fn, poly, xext):
    rows, N = poly.shape[0], poly.shape[1]
    pb = DeviceBuffer.from_array(np.ascontiguousarray(poly, dtype=np.uint64))
    xb = DeviceBuffer.from_array(np.ascontiguousarray(xext, dtype=np.uint64))
    hb = DeviceBuffer(rows * (N // 8) * 144)
    hb.fill(0xA5)
    call(fn, hb.ptr, pb.ptr, xb.ptr, rows, N, None)
    synchronize()
    out = hb.download((rows, N // 8, 18))
    for b in (pb, xb, hb):
        b.free()
    return out


def test_noalloc_equals_poly2h_fft_twice_second_smaller():
    cases = [(4096, 37, 71), (2048, 9, 72)]  # (N, rows, seed): the second call is smaller
    for N, rows, seed in cases:
        print("noalloc: N=%d rows=%d seed=%d" % (N, rows, seed), flush=True)
        tau = S.random_secret(seed)
        poly = S.random_polys(rows, N, seed + 1)
        xext = O.xext_fft_from_tau(tau, N)
        want = run("fk20_poly2h_fft", poly, xext)
        got = run("fk20_poly2h_fft_noalloc", poly, xext)
        assert np.array_equal(got, want)
        # and against the closed form on a few rows
        exp = O.fk20_closed_form_affine(poly[:3], tau)
        bad, first = O.match_affine(got[:3].reshape(-1, 18), exp.reshape(-1, 12))
        assert bad == 0, "row point %d differs from the closed form" % first


def test_noalloc_largest_request_and_range_error():
    N = 4096
    rows = max_rows(N)
    assert rows == 95
    print("noalloc: N=%d rows=%d (largest)" % (N, rows), flush=True)
    tau = S.random_secret(81)
    poly = S.random_polys(rows, N, 82)
    xext = O.xext_fft_from_tau(tau, N)
    assert np.array_equal(run("fk20_poly2h_fft_noalloc", poly, xext), run("fk20_poly2h_fft", poly, xext))
    # one row more does not fit: rejected on the host, nothing launched
    f = lib().fk20_poly2h_fft_noalloc
    assert f(ctypes.c_void_p(16), ctypes.c_void_p(16), ctypes.c_void_p(16), rows + 1, N, None) == FK20_ERANGE
    assert f(ctypes.c_void_p(16), ctypes.c_void_p(16), ctypes.c_void_p(16), max_rows(1024) + 1, 1024, None) == FK20_ERANGE
    assert f(None, None, None, 0, N, None) == 0
