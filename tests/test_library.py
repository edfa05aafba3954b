"""CPU checks of libfk20.so: it loads, exports every function include/fk20.h declares, and
rejects invalid arguments before touching a device."""
import ctypes
import os
import re

import pytest

from fk20_amd import lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "fk20.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|size_t|const char \*)\s*\**\s*(\w+)\s*\(", text, flags=re.M)))


def test_header_parsed():
    names = header_functions()
    assert "fk20_poly2h_fft" in names and "fr_fft" in names and "g1p_ift" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(L.LIBPATH)
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(header_functions()) - {"fk20_status_string"} <= set(L.EXPORTS)


@pytest.mark.parametrize("call", [
    lambda l: l.fk20_poly2toeplitz_coefficients(1, 1, 1, 1000, None),
    lambda l: l.fk20_poly2toeplitz_coefficients(None, 1, 1, 4096, None),
    lambda l: l.fk20_poly2toeplitz_coefficients_fft(1, 1, 1, 512, None),
    lambda l: l.fr_fft(1, 1, 1, 64, None),
    lambda l: l.fr_ift(None, 1, 1, 512, None),
    lambda l: l.g1p_fft(1, 1, 1, 1024, None),
    lambda l: l.g1p_ift(1, None, 2, 512, None),
    lambda l: l.fk20_msm(1, 1, None, 1, 512, None),
    lambda l: l.fk20_msm(1, 1, 1, 1, 100, None),
    lambda l: l.fk20_hext2h(None, 1, 512, None),
    lambda l: l.fk20_hext_fft2h_fft(1, 1, 1, 300, None),
    lambda l: l.fk20_poly2h_fft(1, 1, 1, 1, 4000, None),
    lambda l: l.fk20_poly2h_fft(1, None, 1, 1, 4096, None),
    lambda l: l.fk20_setup2xext_fft(None, 1, 4096, None),
    lambda l: l.fk20_fp_op(9, 1, 1, 1, 1, None),
    lambda l: l.fk20_fr_op(0, 1, 1, None, 1, None),
    lambda l: l.fk20_g1p_op(3, 1, 1, None, None, 1, None),
    lambda l: l.fk20_mad_probe(None, 1, 1, None),
    lambda l: l.fk20_mad_probe(1, 0, 1, None),
    lambda l: l.fk20_fp_mul_probe(30, 16, 16, 16, 1, 1, None),
    lambda l: l.fk20_fp_mul_probe(29, None, 16, 16, 1, 1, None),
    lambda l: l.fk20_fp_mul_probe(32, 16, 16, 16, 0, 1, None),
])
def test_invalid_arguments_rejected(call):
    assert call(L.lib()) == -1  # FK20_EINVAL


def test_zero_rows_is_a_no_op():
    assert L.lib().fk20_poly2toeplitz_coefficients(None, None, 0, 4096, None) == 0
    assert L.lib().fk20_hext2h(None, 0, 512, None) == 0


def test_twiddle_programs_selfcheck():
    """GLV split + wNAF programs of every w^(+-i), i < 512, evaluate to the root over Fr."""
    from fk20_amd import prover
    status, nops, width = prover.twiddle_selfcheck()
    assert status == 0 and width in (3, 4, 5)
    # about 2 x 128 / (width + 1) additions per program (wNAF digits of two 128-bit halves)
    assert 1024 * 0.8 * 256 / (width + 1) < nops < 1024 * 1.3 * 256 / (width + 1)


def test_setup_table_bytes():
    from fk20_amd import prover
    # entries are affine (x, y), 96 bytes, padded to 128
    assert prover.setup_table_bytes(4096, 8) == 16 * 512 * 32 * 128 * 128
    assert prover.setup_table_bytes(1024, 10) == 16 * 128 * 26 * 512 * 128
    assert prover.setup_table_bytes(4096, 13) == 16 * 512 * 20 * 4096 * 128
    assert prover.setup_table_bytes(4096, 7) == 0
    assert prover.setup_table_bytes(4096, 14) == 0


def test_executed_work_model_matches_library_programs():
    """The work model counts the library's own programs (fk20_twiddle_program), the same
    ones the self-check builds, at the library's digit width."""
    from fk20_amd import prover, workload
    _, nops, width = prover.twiddle_selfcheck()
    assert workload.WNAF_W == width
    model = sum(workload.twiddle_program(e, inv)[3] for inv in (False, True) for e in range(512))
    assert model == nops
    # w^128 = z lambda + z: 64-bit halves, about 62 doublings
    dbl, adds, _, _ = workload.twiddle_program(128)
    assert 55 <= dbl <= 66 and adds <= 30
    n = ctypes.c_uint()
    for which, idx, w in ((0, 512, 5), (1, 512, 5), (2, 10, 5), (3, 0, 5), (-1, 0, 5), (0, 1, 2), (0, 1, 6)):
        assert L.lib().fk20_twiddle_program(which, idx, w, None, 0, ctypes.byref(n)) == -1
    # narrower tables mean more additions: wNAF-3 programs are longer than wNAF-5 ones
    lens = [sum(len(prover.twiddle_program(0, e, w)) for e in range(1, 512, 37)) for w in (3, 4, 5)]
    assert lens[0] > lens[1] > lens[2]


def test_schedule_arguments_checked_before_device():
    lib = L.lib()
    for w, r, o in ((2, 4, 2), (6, 4, 2), (5, 3, 2), (5, 4, 1), (5, 4, 4)):
        assert lib.fk20_ctx_set_schedule(None, w, r, o) == -1
    assert lib.fk20_ctx_get_schedule(None, None, None, None) == -1


def test_noalloc_range_check_before_device():
    """fk20_poly2h_fft_noalloc rejects a request larger than its fixed workspace with
    FK20_ERANGE (and bad sizes with FK20_EINVAL) before any device call."""
    f = L.lib().fk20_poly2h_fft_noalloc
    p = ctypes.c_void_p(16)
    assert f(p, p, p, 96, 4096, None) == -6
    assert f(p, p, p, 194, 2048, None) == -6
    assert f(p, p, p, 389, 1024, None) == -6
    assert f(p, p, p, 1, 4095, None) == -1
    assert f(None, p, p, 1, 4096, None) == -1
    assert f(None, None, None, 0, 4096, None) == 0


def test_output_format_sizes_and_argument_checks():
    lib = L.lib()
    assert [lib.fk20_output_point_bytes(f) for f in (0, 1, 2, 3, 4, -1)] == [144, 144, 48, 48, 0, 0]
    assert lib.fk20_ctx_set_output(None, 0) == -1  # FK20_EINVAL
    assert lib.fk20_g1p_compress(None, None, 1, None) == -1
    assert lib.fk20_g1p_compress(None, None, 0, None) == 0


def test_commit_sizes_and_argument_checks():
    lib = L.lib()
    assert lib.fk20_commit_setup_table_bytes(4096, 8) == 4096 * 32 * 128 * 128
    assert lib.fk20_commit_setup_table_bytes(4000, 8) == 0
    assert lib.fk20_commit_workspace_bytes(512, 4096) == (256 + 16) * 512 * 144
    assert lib.fk20_commit_setup_load(None, None, 4096, 8, None, None) == -1
    assert lib.fk20_commit_rows(None, 0, None, None, 1, None) == -1


def test_context_setters_reject_null():
    lib = L.lib()
    assert lib.fk20_ctx_set_streams(None, 2) == -1
    assert lib.fk20_ctx_set_chain(None, 1) == -1
    assert lib.fk20_ctx_set_graph(None, 1) == -1
