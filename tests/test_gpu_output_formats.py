"""GPU output formats of the batched path (fk20_ctx_set_output) and the standalone compressed
serialisation, against the oracle's data-availability order (FK20Py
list_to_reverse_bit_order, kzg_proofs.py:107-118) and compressed encoding."""
import numpy as np
import pytest

import oracle as O
from fk20_amd import ops
from fk20_amd import synthetic as S
from fk20_amd.prover import Prover

pytestmark = pytest.mark.gpu


def test_output_formats_match_oracle():
    N, rows = 1024, 70
    tau = S.random_secret(501)
    poly = S.random_polys(rows, N, 502)
    poly[3] = 0  # a zero row: every proof is the point at infinity
    p = Prover()
    try:
        sid = p.load_setup(O.xext_fft_from_tau(tau, N), N, 8)
        outs = {}
        for fmt in ("g1p", "g1p_da", "compressed", "compressed_da"):
            p.set_output(fmt)
            outs[fmt] = p.prove(sid, poly)
        p.set_output("g1p")
        assert np.array_equal(outs["g1p_da"], O.reverse_bit_order(outs["g1p"]))
        comp = O.g1_compress(outs["g1p"])
        assert comp[3, 0, 0] == 0xC0
        assert np.array_equal(outs["compressed"], comp)
        assert np.array_equal(outs["compressed_da"], O.reverse_bit_order(comp))
        # the natural-order proofs are the reference's, so the formats carry the right points
        G = O.g1p_gen()
        want = np.array([O.g1p_op("mul", G, s) for s in O.fk20_shadow(poly[0], tau)])
        assert np.array_equal(O.to_affine(outs["g1p"][0]), O.to_affine(want))
    finally:
        p.close()


def test_standalone_compress():
    G = O.g1p_gen()
    pts = np.array([G, O.g1p_inf(), O.g1p_op("neg", G)] + [O.g1p_op("mul", G, k) for k in S.random_fr(61, 503)])
    assert np.array_equal(ops.compress(pts), O.g1_compress(pts))
