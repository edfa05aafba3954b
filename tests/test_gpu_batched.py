"""GPU parity of the batched path (context API: fk20_setup_load + fk20_prove_rows) against
the oracle, the scalar-shadow closed form, and the plain per-row path."""
import numpy as np
import pytest

import oracle as O
from fk20_amd import ops
from fk20_amd import synthetic as S
from fk20_amd.prover import Prover

pytestmark = pytest.mark.gpu


def aff(pts):
    return O.to_affine(np.ascontiguousarray(pts, dtype=np.uint64).reshape(-1, 18))


def shadow_points(poly, tau):
    G = O.g1p_gen()
    return np.array([O.g1p_op("mul", G, s) for s in O.fk20_shadow(poly, tau)])


@pytest.fixture(scope="module")
def prover():
    p = Prover()
    yield p
    p.close()


@pytest.mark.parametrize("N,rows,window", [(1024, 70, 8), (2048, 65, 10)])
def test_batched_matches_plain_and_closed_form(prover, N, rows, window):
    tau = S.random_secret(N + window)
    poly = S.random_polys(rows, N, N + 7)
    xext = O.xext_fft_from_tau(tau, N)
    sid = prover.load_setup(xext, N, window)
    got = prover.prove(sid, poly)
    plain = ops.poly2h_fft(poly, xext)
    assert np.array_equal(aff(got), aff(plain))
    for r in (0, 63, 64, rows - 1):
        assert np.array_equal(aff(got[r]), aff(shadow_points(poly[r], tau))), r
    prover.unload_setup(sid)


def test_batched_4096(prover):
    N, rows = 4096, 64
    tau = S.random_secret(99)
    poly = S.random_polys(rows, N, 98)
    xext = O.xext_fft_from_tau(tau, N)
    sid = prover.load_setup(xext, N, 8)
    got = prover.prove(sid, poly)
    plain = ops.poly2h_fft(poly[:8], xext)
    assert np.array_equal(aff(got[:8]), aff(plain))
    for r in (0, 37, 63):
        assert np.array_equal(aff(got[r]), aff(shadow_points(poly[r], tau))), r
    prover.unload_setup(sid)


def test_batched_edge_inputs(prover):
    """Zero rows, a setup point at infinity, and non-normalised projective setup points."""
    N, rows = 1024, 3
    tau = S.random_secret(5)
    xext = O.xext_fft_from_tau(tau, N).copy()
    xext[3, 17] = O.g1p_inf()
    # scale one point's coordinates by 5 (same projective point)
    for c in range(3):
        v = O.limbs_to_int(xext[7, 9, 6 * c: 6 * c + 6]) * 5 % O.P_MOD
        xext[7, 9, 6 * c: 6 * c + 6] = O.fp(v)
    poly = S.random_polys(rows, N, 6)
    poly[1] = 0
    sid = prover.load_setup(xext, N, 8)
    got = prover.prove(sid, poly)
    exp = O.poly2h_fft(poly, xext)
    assert np.array_equal(aff(got), aff(exp))
    assert all(O.g1p_is_inf(p) for p in got[1])
    prover.unload_setup(sid)


def test_batched_msm_exceptional_sums(prover):
    """Column sums of the batched MSM that add a point to itself or to its negative: every
    setup row holds the same point Q_j (row 1 holds -Q_j for odd j), and polynomials whose
    top 16 coefficients are c (the rest 0) have Toeplitz coefficients c at position 0 only,
    so every MSM scalar is c and the second term meets the first one's sum.  Against the
    plain path and the oracle; 20 rows, so the row-parallel MSM kernel runs."""
    N, rows = 1024, 20
    n2 = N // 8
    q = ops.g1p_op("mul", np.tile(O.g1p_gen(), (n2, 1)), k=S.random_fr(n2, 570))
    xext = np.ascontiguousarray(np.repeat(q[None], 16, axis=0))
    for j in range(1, n2, 2):
        y = O.limbs_to_int(xext[1, j, 6:12])
        xext[1, j, 6:12] = O.fp((O.P_MOD - y) % O.P_MOD)
    poly = S.random_polys(rows, N, 571)
    for r, c in ((0, 1), (2, 2), (5, 4095)):
        poly[r] = 0
        poly[r, N - 16:] = O.fr(c)
    poly[3] = 0
    sid = prover.load_setup(xext, N, 8)
    got = prover.prove(sid, poly)
    assert np.array_equal(aff(got), aff(ops.poly2h_fft(poly, xext)))
    for r in (0, 2, 5):
        assert np.array_equal(aff(got[r]), aff(O.poly2h_fft(poly[r:r + 1], xext))), r
    assert all(O.g1p_is_inf(p) for p in got[3])
    prover.unload_setup(sid)


def test_graph_replay_tracks_input_contents():
    from fk20_amd.lib import DeviceBuffer, Stream
    N, rows = 1024, 64
    tau = S.random_secret(8)
    xext = O.xext_fft_from_tau(tau, N)
    p = Prover(graph=True)
    try:
        sid = p.load_setup(xext, N, 8)
        st = Stream()
        a = S.random_polys(rows, N, 1)
        b = S.random_polys(rows, N, 2)
        pb = DeviceBuffer.from_array(a)
        hb = DeviceBuffer(rows * (N // 8) * 144)
        outs = []
        for src in (a, b, a):
            pb.upload(src)
            p.prove_device(sid, hb, pb, rows, stream=st.handle)
            st.synchronize()
            outs.append(hb.download((rows, N // 8, 18)))
        ref_a = p.prove(sid, a)
        ref_b = p.prove(sid, b)
        assert np.array_equal(aff(outs[0]), aff(ref_a))
        assert np.array_equal(aff(outs[1]), aff(ref_b))
        assert np.array_equal(aff(outs[2]), aff(ref_a))
        st.destroy()
    finally:
        p.close()


def test_uneven_batch_one_pass(prover):
    """Rows of N = 1024 / 2048 / 4096 with three setups in one fk20_prove_batch call."""
    groups, taus = [], {}
    for N, rows, seed in ((1024, 70, 301), (2048, 33, 302), (4096, 20, 303)):
        tau = S.random_secret(seed)
        xext = O.xext_fft_from_tau(tau, N)
        sid = prover.load_setup(xext, N, 8)
        poly = S.random_polys(rows, N, seed + 10)
        groups.append((sid, poly))
        taus[sid] = (tau, xext)
    outs = prover.prove_batch(groups)
    for (sid, poly), got in zip(groups, outs):
        tau, xext = taus[sid]
        alone = prover.prove(sid, poly)
        assert np.array_equal(aff(got), aff(alone))
        for r in (0, poly.shape[0] - 1):
            assert np.array_equal(aff(got[r]), aff(shadow_points(poly[r], tau)))
    for sid in list(taus):
        prover.unload_setup(sid)


def test_uneven_small_batch_one_pass(prover):
    """Items of at most 16 rows (the MSM's term-parallel kernel) across three sizes, against
    the closed form row by row and a large batch of the same rows."""
    groups, taus = [], {}
    for N, rows, seed in ((1024, 1, 311), (2048, 7, 312), (4096, 16, 313)):
        tau = S.random_secret(seed)
        sid = prover.load_setup(O.xext_fft_from_tau(tau, N), N, 8)
        groups.append((sid, S.random_polys(rows, N, seed + 10)))
        taus[sid] = tau
    outs = prover.prove_batch(groups)
    for (sid, poly), got in zip(groups, outs):
        for r in range(poly.shape[0]):
            assert np.array_equal(aff(got[r]), aff(shadow_points(poly[r], taus[sid]))), (sid, r)
    # the 16-row group inside a 17-row batch takes the row-parallel kernel: same points
    sid, poly = groups[2]
    big = prover.prove(sid, np.concatenate([poly, poly[:1]]))
    assert np.array_equal(aff(big[:16]), aff(outs[2]))
    for sid in list(taus):
        prover.unload_setup(sid)


def test_chain_equals_per_stage_launches(prover):
    """The persistent transform chain (radix-2 schedule) and one launch per stage (default,
    radix-2^2 schedule) give the same points (canonical affine; the projective representatives
    differ with the schedule), on an uneven batch with partial 64-row groups; the chain also
    matches the closed form."""
    groups, taus = [], []
    for N, rows, seed in ((1024, 70, 401), (4096, 66, 402), (2048, 1, 403)):
        tau = S.random_secret(seed)
        sid = prover.load_setup(O.xext_fft_from_tau(tau, N), N, 8)
        groups.append((sid, S.random_polys(rows, N, seed + 10)))
        taus.append(tau)
    staged = prover.prove_batch(groups)
    try:
        prover.set_chain(True)
        chained = prover.prove_batch(groups)
    finally:
        prover.set_chain(False)
    for (sid, poly), tau, a, b in zip(groups, taus, chained, staged):
        assert np.array_equal(aff(a), aff(b))
        r = poly.shape[0] - 1
        assert np.array_equal(aff(a[r]), aff(shadow_points(poly[r], tau)))
    for sid, _ in groups:
        prover.unload_setup(sid)


def test_every_schedule_variant_gives_the_same_points(prover):
    """Each transform schedule the autotuner may pick -- wNAF width 3 / 4 / 5, radix-2^2 or
    radix-2 butterflies, two or three waves per SIMD, and the persistent chain at each width --
    proves an uneven batch (partial 64-row groups, three sizes) to the same points as the
    default schedule (canonical affine), and the default matches the closed form."""
    groups, taus = [], []
    for N, rows, seed in ((1024, 70, 411), (4096, 3, 412), (2048, 65, 413)):
        tau = S.random_secret(seed)
        sid = prover.load_setup(O.xext_fft_from_tau(tau, N), N, 8)
        groups.append((sid, S.random_polys(rows, N, seed + 10)))
        taus.append(tau)
    base = [aff(x) for x in prover.prove_batch(groups)]
    for (sid, poly), tau, b in zip(groups, taus, base):
        r = poly.shape[0] - 1
        assert np.array_equal(b[r * (b.shape[0] // poly.shape[0]):], aff(shadow_points(poly[r], tau)))
    try:
        for wnaf in (3, 4, 5):
            for radix in (2, 4):
                for waves in (2, 3):
                    prover.set_schedule(wnaf, radix, waves)
                    assert prover.schedule() == {"wnaf": wnaf, "radix": radix, "waves_per_simd": waves}
                    got = prover.prove_batch(groups)
                    for a, b in zip(got, base):
                        assert np.array_equal(aff(a), b), (wnaf, radix, waves)
            prover.set_schedule(wnaf, 4, 2)
            prover.set_chain(True)
            try:
                got = prover.prove_batch(groups)
            finally:
                prover.set_chain(False)
            for a, b in zip(got, base):
                assert np.array_equal(aff(a), b), (wnaf, "chain")
    finally:
        prover.set_schedule(5, 4, 2)
    for sid, _ in groups:
        prover.unload_setup(sid)


def test_two_stream_split_equals_single_stream(prover):
    """The default two-halves launch layout gives bit-identical outputs to one launch sequence,
    for items of one, two and three 64-row groups in one batch, directly and from a graph."""
    groups = []
    for N, rows, seed in ((2048, 40, 601), (1024, 128, 602), (4096, 130, 603)):
        sid = prover.load_setup(O.xext_fft_from_tau(S.random_secret(seed), N), N, 8)
        groups.append((sid, S.random_polys(rows, N, seed + 10)))
    split = prover.prove_batch(groups)
    try:
        prover.set_streams(1)
        single = prover.prove_batch(groups)
    finally:
        prover.set_streams(2)
    for a, b in zip(split, single):
        assert np.array_equal(a, b)
    for sid, _ in groups:
        prover.unload_setup(sid)


def test_chain_expired_wait_is_an_error(prover):
    """A zero wait budget makes every wait of the persistent launch that finds its inputs
    unfinished give up at once (each wave then runs its butterfly and exits normally); the
    call must report FK20_ETIMEDOUT instead of returning the stale proofs as valid.  With the
    default budget restored the same context proves correctly again."""
    from fk20_amd.lib import Fk20Error
    N, rows = 1024, 128
    tau = S.random_secret(411)
    sid = prover.load_setup(O.xext_fft_from_tau(tau, N), N, 8)
    poly = S.random_polys(rows, N, 412)
    try:
        prover.set_chain(True)
        prover.set_chain_wait(0)
        with pytest.raises(Fk20Error) as ei:
            prover.prove(sid, poly)
        assert ei.value.status == -5
        prover.set_chain_wait(1 << 27)
        got = prover.prove(sid, poly)
    finally:
        prover.set_chain_wait(1 << 27)
        prover.set_chain(False)
    for r in (0, rows - 1):
        assert np.array_equal(aff(got[r]), aff(shadow_points(poly[r], tau)))
    prover.unload_setup(sid)


def test_calls_on_two_streams_share_the_workspace_in_order(prover):
    """Two fk20_prove_rows calls on different streams, enqueued back to back with no host
    synchronisation between them, use the context's one workspace: the second must wait for
    the first on the device.  Small batches (one 64-row group: no internal split) are the case
    where both run on the callers' streams directly."""
    from fk20_amd.lib import DeviceBuffer, Stream
    N, rows = 1024, 48
    tau = S.random_secret(421)
    sid = prover.load_setup(O.xext_fft_from_tau(tau, N), N, 8)
    polys = [S.random_polys(rows, N, 422 + k) for k in range(3)]
    streams = [Stream(), Stream()]
    pbs = [DeviceBuffer.from_array(p) for p in polys]
    hbs = [DeviceBuffer(rows * (N // 8) * 144) for _ in polys]
    for k in range(3):
        prover.prove_device(sid, hbs[k], pbs[k], rows, stream=streams[k % 2].handle)
    for st in streams:
        st.synchronize()
    outs = [hb.download((rows, N // 8, 18)) for hb in hbs]
    for b in pbs + hbs:
        b.free()
    for st in streams:
        st.destroy()
    for poly, got in zip(polys, outs):
        assert O.match_affine(got, O.fk20_closed_form_affine(poly, tau)) == (0, -1)
    prover.unload_setup(sid)
