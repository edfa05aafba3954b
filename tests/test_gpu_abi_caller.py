"""A compiled C++ caller of include/fk20.h (tests/native/abi_caller.cpp, linked with -lfk20,
built by __graft_entry__.build()) runs the host sequence a reference caller switches to:
fk20_poly2h_fft, fk20_g1p_to_affine, fk20_poly2h_fft_noalloc and the context path
(fk20_ctx_create, fk20_setup_load, fk20_ctx_reserve, fk20_prove_rows on a stream).  Its
outputs are checked against the oracle's closed form h_fft[k] = [s_k]G (canonical affine,
bit-exact) and against each other."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from fk20_amd import synthetic as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "abi_caller.cpp")
BIN = os.path.join(ROOT, "tests", "native", "bin", "abi_caller")


def build_caller():
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", BIN, SRC,
                    "-L", os.path.join(ROOT, "fk20-cuda_amd"), "-lfk20",
                    "-Wl,-rpath,$ORIGIN/../../../fk20-cuda_amd"], check=True)


@pytest.mark.parametrize("N,rows,seed", [(1024, 3, 501), (4096, 2, 502)])
def test_compiled_caller_matches_oracle(tmp_path, N, rows, seed):
    if not os.path.exists(BIN):
        build_caller()
    tau = S.random_secret(seed)
    poly = np.ascontiguousarray(S.random_polys(rows, N, seed + 1), dtype=np.uint64)
    xext = np.ascontiguousarray(O.xext_fft_from_tau(tau, N), dtype=np.uint64)
    poly.tofile(tmp_path / "poly.bin")
    xext.tofile(tmp_path / "xext.bin")
    print("abi_caller: N=%d rows=%d" % (N, rows), flush=True)
    r = subprocess.run([BIN, str(tmp_path), str(rows), str(N)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr + r.stdout
    n2 = N // 8
    plain = np.fromfile(tmp_path / "h_plain.bin", dtype=np.uint64).reshape(rows * n2, 18)
    noalloc = np.fromfile(tmp_path / "h_noalloc.bin", dtype=np.uint64).reshape(rows * n2, 18)
    ctx = np.fromfile(tmp_path / "h_ctx.bin", dtype=np.uint64).reshape(rows * n2, 18)
    aff = np.fromfile(tmp_path / "h_affine.bin", dtype=np.uint64).reshape(rows * n2, 12)
    exp = O.fk20_closed_form_affine(poly, tau).reshape(-1, 12)
    assert np.array_equal(aff, exp)
    for name, got in (("fk20_poly2h_fft", plain), ("fk20_prove_rows", ctx)):
        bad, first = O.match_affine(got, exp)
        assert bad == 0, "%s: point %d differs" % (name, first)
    assert np.array_equal(noalloc, plain)
