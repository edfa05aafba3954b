"""Serving loop (fk20_server_*): each request's result is identical whether it is proved alone
or admitted into continuous batches mixed with other sizes and setups."""
import threading
import time

import numpy as np
import pytest

import oracle as O
from fk20_amd import synthetic as S
from fk20_amd.lib import Fk20Error, PinnedArray
from fk20_amd.prover import Prover
from fk20_amd.server import Server

pytestmark = pytest.mark.gpu


def aff(pts):
    return O.to_affine(np.ascontiguousarray(pts, dtype=np.uint64).reshape(-1, 18))


def test_mix_equals_alone():
    p = Prover()
    try:
        sids = {}
        for N, seed in ((1024, 501), (2048, 502), (4096, 503)):
            sids[N] = p.load_setup(O.xext_fft_from_tau(S.random_secret(seed), N), N, 8)
        rng = np.random.default_rng(7)
        reqs = []
        for k in range(36):
            N = (1024, 2048, 4096, 4096)[k % 4]
            rows = int(rng.integers(1, 4))
            poly = PinnedArray((rows, N, 4))
            poly.array[:] = S.random_polys(rows, N, 1000 + k)
            out = PinnedArray((rows, N // 8, 18))
            out.array[:] = 0xAA  # canary
            reqs.append((N, rows, poly, out))
        srv = Server(p, budget_bytes=256 << 20, slots=2, max_batch_rows=24)
        ids = []

        def submit(part):
            for N, rows, poly, out in part:
                ids.append(srv.submit(sids[N], poly.ptr, rows, out.ptr))
                time.sleep(0.002)

        # arrivals in three waves while earlier batches are on the device
        for w in range(3):
            th = threading.Thread(target=submit, args=(reqs[w::3],))
            th.start()
            th.join()
        lat = [srv.wait(i) for i in ids]
        st = srv.stats()
        srv.close()
        assert st["status"] == 0 and st["rows_done"] == sum(r[1] for r in reqs)
        assert st["batches"] >= 2 and all(x > 0 for x in lat)
        for N, rows, poly, out in reqs:
            alone = p.prove(sids[N], poly.array)
            assert np.array_equal(out.array, alone)  # bit-identical projective output
        # one request against the closed form (independent of the batched path)
        N, rows, poly, out = reqs[2]
        tau = S.random_secret({1024: 501, 2048: 502, 4096: 503}[N])
        G = O.g1p_gen()
        exp = np.array([O.g1p_op("mul", G, s) for s in O.fk20_shadow(poly.array[0], tau)])
        assert np.array_equal(aff(out.array[0]), aff(exp))
        for _, _, poly, out in reqs:
            poly.free()
            out.free()
    finally:
        p.close()


def test_request_over_budget_is_rejected():
    p = Prover()
    try:
        sid = p.load_setup(O.xext_fft_from_tau(S.random_secret(9), 1024), 1024, 8)
        srv = Server(p, budget_bytes=64 << 20, slots=2, max_batch_rows=4096)
        poly = PinnedArray((1000, 1024, 4))
        out = PinnedArray((1000, 128, 18))
        with pytest.raises(Fk20Error) as ei:
            srv.submit(sid, poly.ptr, 1000, out.ptr)  # 1000 rows x ~148 KB > the 64 MB budget
        assert ei.value.status == -2
        srv.close()
        poly.free()
        out.free()
    finally:
        p.close()


def test_budget_is_shared_by_the_slots():
    """One arena serves every slot: a request larger than half the budget (the former per-slot
    share) is admitted whole and proved like the same rows alone; two of them queue behind
    each other instead of failing."""
    p = Prover()
    try:
        N = 1024
        tau = S.random_secret(31)
        sid = p.load_setup(O.xext_fft_from_tau(tau, N), N, 8)
        rows = 300  # ~44 MB at N = 1024 (about 148 KB per row), within 64 MB, above 32 MB
        poly = PinnedArray((2, rows, N, 4))
        poly.array[:] = S.random_polys(2 * rows, N, 32).reshape(2, rows, N, 4)
        out = PinnedArray((2, rows, N // 8, 18))
        srv = Server(p, budget_bytes=64 << 20, slots=2, max_batch_rows=4096)
        ids = [srv.submit(sid, poly.ptr + k * rows * N * 32, rows, out.ptr + k * rows * (N // 8) * 144)
               for k in range(2)]
        for i in ids:
            srv.wait(i)
        st = srv.stats()
        srv.close()
        assert st["batches"] == 2 and st["max_inflight_rows"] == rows
        alone = p.prove(sid, poly.array[1])
        assert np.array_equal(out.array[1], alone)
        exp = O.fk20_closed_form_affine(poly.array[0][:2], tau)
        bad, first = O.match_affine(out.array[0][:2].reshape(-1, 18), exp.reshape(-1, 12))
        assert bad == 0
        poly.free()
        out.free()
    finally:
        p.close()
