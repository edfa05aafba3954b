"""End-to-end validation of GPU output by the KZG multi-proof check (FK20Py check_proof_multi,
kzg_proofs.py:207, with oracle/pairing.py's pairing): proofs from the batched path and
commitments from fk20_commit_rows, at the benchmark size N = 4096, verify against the true
evaluations on each proof's coset; a proof of another coset does not."""
import numpy as np
import pytest

import oracle as O
import pairing as PR
from fk20_amd import synthetic as S
from fk20_amd.prover import Prover

pytestmark = pytest.mark.gpu


def test_gpu_proofs_and_commitments_verify():
    N, rows = 4096, 64
    tau = S.random_secret(801)
    poly = S.random_polys(rows, N, 802)
    srs = O.setup_srs(tau, N)
    p = Prover()
    try:
        fid = p.load_setup(O.xext_fft_from_tau(tau, N), N, 8)
        cid = p.load_commit_setup(srs, N, 8)
        proofs = p.prove(fid, poly)
        commits = p.commit(cid, poly)
    finally:
        p.close()
    srs_aff = [O.g1_affine(srs[i]) for i in range(16)]
    g2_s16 = PR.g2_mul(PR.G2_GEN, pow(tau, 16, PR.R))
    w16 = pow(5, (PR.R - 1) // 16, PR.R)
    for row, k in ((0, 0), (37, 511)):
        x = O.fk20_proof_coset(k, N)
        ys = [O.poly_eval(poly[row], x * pow(w16, i, PR.R) % PR.R) for i in range(16)]
        c = O.g1_affine(commits[row])
        assert PR.check_proof_multi(c, O.g1_affine(proofs[row, k]), x, ys, srs_aff, g2_s16,
                                    O.g1_affine_add, O.g1_affine_mul)
        assert not PR.check_proof_multi(c, O.g1_affine(proofs[row, (k + 1) % 512]), x, ys, srs_aff, g2_s16,
                                        O.g1_affine_add, O.g1_affine_mul)
