"""The benchmarked configurations, every output point against the oracle's closed form.

C2: 512 rows of N = 4096 on one setup with the MSM window tables the autotuner picks (13
bits) and the next size down (12 bits): fk20_prove_rows with the rows split over two
streams (default), one launch sequence, hipGraph replay, and batches of at most 16 rows
(the term-parallel MSM kernel).  C4: the uneven batch, 512 rows each of N = 1024 / 2048 /
4096 on three setups in one fk20_prove_batch call.  C3: the drop-in entry points at their
microbenchmark shapes -- fk20_msm over 512 rows x 512 columns, 512 G1 FFTs and IFFTs of 512
points, 8192 NTTs of 512 -- also on every output.

Expected values come from the scalar shadow: the setups are built from a known secret, so
h_fft[k] = [s_k]G with s_k the FK20 algorithm run over Fr (oracle or_fk20_shadow, pinned
against KZG quotient openings in test_oracle.py), and [s]G is computed by fixed-base windows
(or_g1_gen_mul_affine_n, pinned against double-and-add).  The reference checks its 512-row
batch the same way, every row and stage against independently generated data
(fk20_512test.cu:77-118).
"""
import numpy as np
import pytest

import oracle as O
from fk20_amd import ops
from fk20_amd import synthetic as S
from fk20_amd.lib import DeviceBuffer, Stream, synchronize
from fk20_amd.prover import Prover

pytestmark = pytest.mark.gpu


def assert_all_match(got, exp_aff, what):
    bad, first = O.match_affine(np.ascontiguousarray(got, dtype=np.uint64).reshape(-1, 18),
                                np.ascontiguousarray(exp_aff).reshape(-1, 12))
    assert bad == 0, "%s: %d points differ, first at flat index %d" % (what, bad, first)


@pytest.fixture(scope="module")
def c2():
    N, rows = 4096, 512
    tau = S.random_secret(2025)
    poly = S.random_polys(rows, N, 2024)
    xext = O.xext_fft_from_tau(tau, N)
    exp = O.fk20_closed_form_affine(poly, tau)
    return N, rows, tau, poly, xext, exp


@pytest.fixture(scope="module")
def w13(c2):
    N, rows, tau, poly, xext, exp = c2
    p = Prover()
    sid = p.load_setup(xext, N, 13)
    yield p, sid
    p.close()


def test_c2_w13_two_streams_every_row(c2, w13):
    N, rows, tau, poly, xext, exp = c2
    p, sid = w13
    p.set_streams(2)
    assert_all_match(p.prove(sid, poly), exp, "w13 two streams")


def test_c2_w13_one_stream_every_row(c2, w13):
    N, rows, tau, poly, xext, exp = c2
    p, sid = w13
    p.set_streams(1)
    try:
        assert_all_match(p.prove(sid, poly), exp, "w13 one stream")
    finally:
        p.set_streams(2)


def test_c2_w13_graph_replay_every_row(c2, w13):
    N, rows, tau, poly, xext, exp = c2
    p, sid = w13
    st = Stream()
    pb = DeviceBuffer.from_array(poly)
    hb = DeviceBuffer(rows * (N // 8) * 144)
    p.set_graph(True)
    try:
        for rep in range(3):  # capture, then two replays; eager work between them
            hb.fill(0)
            p.prove_device(sid, hb, pb, rows, stream=st.handle)
            st.synchronize()
            assert_all_match(hb.download((rows, N // 8, 18)), exp, "w13 graph replay %d" % rep)
    finally:
        p.set_graph(False)
        st.destroy()
        pb.free()
        hb.free()


@pytest.mark.parametrize("nrows", [1, 5, 16, 17, 64, 65])
def test_c2_w13_small_batches(c2, w13, nrows):
    """<= 16 rows run the term-parallel MSM kernel (k_msm_multi_small); 17, 64, 65 the
    row-parallel one with partial 64-row groups, on the same resident w = 13 tables."""
    N, rows, tau, poly, xext, exp = c2
    p, sid = w13
    lo = 300
    assert_all_match(p.prove(sid, poly[lo:lo + nrows]), exp[lo:lo + nrows], "w13 %d rows" % nrows)


def test_c2_w12_every_row(c2):
    N, rows, tau, poly, xext, exp = c2
    with Prover() as p:
        sid = p.load_setup(xext, N, 12)
        assert_all_match(p.prove(sid, poly), exp, "w12 two streams")


def test_c4_uneven_batch_w13_every_row():
    """512 rows each of N = 1024 / 2048 / 4096, three setups (w = 13: 112 GiB of tables), one
    fk20_prove_batch call; every row of every group against its own setup's closed form."""
    groups, exps = [], []
    with Prover() as p:
        for k, N in enumerate((1024, 2048, 4096)):
            tau = S.random_secret(3100 + k)
            poly = S.random_polys(512, N, 3200 + k)
            sid = p.load_setup(O.xext_fft_from_tau(tau, N), N, 13)
            groups.append((sid, poly))
            exps.append(O.fk20_closed_form_affine(poly, tau))
        outs = p.prove_batch(groups)
    for (sid, poly), got, exp in zip(groups, outs, exps):
        assert_all_match(got, exp, "uneven group N=%d" % poly.shape[1])


# ---------------------------------------------------------------- C3 drop-in shapes

def random_points(scalars, seed):
    """[k]G for each scalar as a non-normalised projective representative (random z)."""
    zs = S.random_fp_nonzero(scalars.size // 4, seed).reshape(scalars.shape[:-1] + (6,))
    return O.gen_mul_projective(scalars, zs)


def test_c3_fk20_msm_512_rows_512_columns():
    rows, n2 = 512, 512
    tc = S.random_fr(rows * 16 * n2, 41).reshape(rows, 16, n2, 4)
    ks = S.random_fr(16 * n2, 42).reshape(16, n2, 4)
    xext = random_points(ks, 43)
    got = ops.msm(tc, xext)
    assert_all_match(got, O.gen_mul_affine(O.msm_shadow(tc, ks)), "fk20_msm 512 x 512")


@pytest.mark.parametrize("inverse", [False, True])
def test_c3_g1p_fft_512_transforms(inverse):
    count, n = 512, 512
    ks = S.random_fr(count * n, 51 + inverse).reshape(count, n, 4)
    pts = random_points(ks, 53 + inverse)
    got = ops.g1p_fft(pts, inverse=inverse)
    exp_scalars = np.stack([O.fr_fft(ks[t], inverse=inverse) for t in range(count)])
    assert_all_match(got, O.gen_mul_affine(exp_scalars), "g1p_%s x512" % ("ift" if inverse else "fft"))


@pytest.mark.parametrize("inverse", [False, True])
def test_c3_fr_fft_8192_transforms(inverse):
    count, n = 8192, 512
    x = S.random_fr(count * n, 61 + inverse).reshape(count, n, 4)
    got = ops.fr_fft(x, inverse=inverse)
    for t in range(count):
        if not np.array_equal(got[t], O.fr_fft(x[t], inverse=inverse)):
            raise AssertionError("fr_%s transform %d differs" % ("ift" if inverse else "fft", t))
