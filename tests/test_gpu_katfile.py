"""Known-answer files at the reference's size (512 rows, N = 4096, fk20_512_kat.bin layout):
bootstrap one from the GPU stage outputs, pin rows of it against the oracle, then run the stage
tests of fk20_512test.cu against it -- plainly, where every stage must match, and with mangled
inputs, where every stage must differ."""
import numpy as np
import pytest

import oracle as O
from fk20_amd import katfile as K

pytestmark = pytest.mark.gpu


def aff(pts):
    return O.to_affine(np.ascontiguousarray(pts, dtype=np.uint64).reshape(-1, 18))


@pytest.fixture(scope="module")
def kat(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("kat") / "fk20_512_kat.bin")
    nbytes, tau = K.bootstrap(path, rows=512, N=4096, seed=512)
    return path, nbytes, tau


def test_bootstrap_size_and_refusal(kat):
    path, nbytes, tau = kat
    assert nbytes == 449970176
    with pytest.raises(FileExistsError):
        K.bootstrap(path, rows=1, N=4096)
    assert K.rows_of(449970176) == 512 and len(K.read(path)["h_fft"]) == 512


def test_file_rows_against_oracle(kat):
    """Row 0 at every stage and row 511's proofs, from the oracle and the closed form."""
    path, _, tau = kat
    k = K.read(path)
    N = 4096
    xext = O.xext_fft_from_tau(tau, N)
    assert np.array_equal(aff(k["xext_fft"]), aff(xext))
    poly0 = np.asarray(k["polynomial"][0])
    tc = O.poly2toeplitz_coefficients(poly0)
    assert np.array_equal(np.asarray(k["toeplitz_coefficients"][0]), tc)
    tcf = np.stack([O.fr_fft(tc[i]) for i in range(16)])
    got_tcf = np.asarray(k["toeplitz_coefficients_fft"][0]).reshape(-1, 4)
    assert all(O.limbs_to_int(a) % O.R_MOD == O.limbs_to_int(b) % O.R_MOD
               for a, b in zip(got_tcf, tcf.reshape(-1, 4)))
    hext = O.msm(tcf, xext)
    assert np.array_equal(aff(k["hext_fft"][0]), aff(hext))
    h = O.g1p_fft(hext, inverse=True)
    h[256:] = O.g1p_inf()
    assert np.array_equal(aff(k["h"][0]), aff(h))
    G = O.g1p_gen()
    for r in (0, 511):
        exp = np.array([O.g1p_op("mul", G, s) for s in O.fk20_shadow(np.asarray(k["polynomial"][r]), tau)])
        assert np.array_equal(aff(k["h_fft"][r]), aff(exp)), r


@pytest.mark.parametrize("rows", [2, 512])
def test_stage_tests_pass(kat, rows):
    report = K.verify(kat[0], rows=rows)
    assert [r["stage"] for r in report] == list(K.STAGE_NAMES)
    bad = [r for r in report if not r["pass"]]
    assert not bad, bad


def test_falsifiability(kat):
    report = K.verify(kat[0], rows=512, falsify=True)
    bad = [r for r in report if not r["pass"]]
    assert not bad, bad
