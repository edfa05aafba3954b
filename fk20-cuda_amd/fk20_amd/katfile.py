"""Batch known-answer files in the layout of FK20-CUDA's `fk20_512_kat.bin`
(fk20_512test_bootstrap.cu:44-59, read back at :103-150): seven headerless sections of
reference-ABI arrays (fr_t = 4 little-endian uint64 limbs, g1p_t = 18), in this order:

    xext_fft                    g1p[16][n2]        the setup (shared by every row)
    polynomial                  fr[rows][N]
    toeplitz_coefficients       fr[rows][16][n2]
    toeplitz_coefficients_fft   fr[rows][16][n2]
    hext_fft                    g1p[rows][n2]
    h                           g1p[rows][n2]
    h_fft                       g1p[rows][n2]

with n2 = N / 8.  The reference fixes rows = 512 and N = 4096 (449 970 176 bytes); here both
are parameters and a file's row count follows from its size.

`bootstrap` writes such a file from the drop-in entry points' stage outputs on the GPU (the
role of fk20_512test_bootstrap.cu); `verify` runs fk20_512test.cu's sequence of stage tests
against a file -- each stage from the file's own input, so that one failing stage does not
cascade -- and `mangle` is its varMangle (:961-1004) for the falsifiability runs, in which
every stage must report mismatches.  Points are compared as projective points (cross
multiplication, the reference's g1p_eq) and field elements modulo r (fr_eq), on the GPU.
"""
import os
import time

import numpy as np

from . import lib as L
from . import synthetic as S

SECTIONS = ("xext_fft", "polynomial", "toeplitz_coefficients", "toeplitz_coefficients_fft",
            "hext_fft", "h", "h_fft")
REFERENCE_ROWS, REFERENCE_N = 512, 4096
MANGLE_START = 3  # fk20_512test.cu:961 START_INDEX


def shapes(rows, N):
    """uint64 array shape of every section."""
    n2 = N // 8
    return {"xext_fft": (16, n2, 18), "polynomial": (rows, N, 4),
            "toeplitz_coefficients": (rows, 16, n2, 4), "toeplitz_coefficients_fft": (rows, 16, n2, 4),
            "hext_fft": (rows, n2, 18), "h": (rows, n2, 18), "h_fft": (rows, n2, 18)}


def file_bytes(rows, N=REFERENCE_N):
    return sum(8 * int(np.prod(s)) for s in shapes(rows, N).values())


def rows_of(nbytes, N=REFERENCE_N):
    """Row count of a file of nbytes bytes, or ValueError if no row count gives that size."""
    fixed = file_bytes(0, N)
    per_row = file_bytes(1, N) - fixed
    if nbytes < fixed + per_row or (nbytes - fixed) % per_row:
        raise ValueError("%d bytes is not a known-answer file for N=%d" % (nbytes, N))
    return (nbytes - fixed) // per_row


def read(path, N=REFERENCE_N, rows=None):
    """Memory-mapped, read-only views of every section, keyed by name.  rows, when given, must
    match the file's size (the reference reads exactly 449 970 176 bytes, :124-150)."""
    nbytes = os.path.getsize(path)
    r = rows_of(nbytes, N)
    if rows is not None and rows != r:
        raise ValueError("%s holds %d rows, not %d" % (path, r, rows))
    out, off = {}, 0
    for name, shp in shapes(r, N).items():
        out[name] = np.memmap(path, dtype=np.uint64, mode="r", offset=off, shape=shp)
        off += 8 * int(np.prod(shp))
    return out


def write(path, arrays, overwrite=False):
    """Write the seven sections.  Like the reference writer (:27-31), an existing file is not
    replaced unless overwrite is set (FileExistsError)."""
    xext = np.asarray(arrays["xext_fft"])
    rows, N = np.asarray(arrays["polynomial"]).shape[:2]
    if xext.shape != (16, N // 8, 18):
        raise ValueError("xext_fft shape %s does not match N=%d" % (xext.shape, N))
    want = shapes(rows, N)
    for name in SECTIONS:
        if tuple(np.shape(arrays[name])) != want[name]:
            raise ValueError("%s has shape %s, expected %s" % (name, np.shape(arrays[name]), want[name]))
    with open(path, "wb" if overwrite else "xb") as f:
        for name in SECTIONS:
            f.write(np.ascontiguousarray(arrays[name], dtype=np.uint64).tobytes())
    return file_bytes(rows, N)


def mangle(a, step, start=MANGLE_START):
    """A copy of a with elements i and i + 1 swapped for i = start, start + step, ... while
    i + step < len (fk20_512test.cu:971-1004, over the flattened element sequence).
    Applying it twice restores the input."""
    a = np.array(a, dtype=np.uint64, copy=True)
    width = a.shape[-1]
    flat = a.reshape(-1, width)
    idx = np.arange(start, flat.shape[0], step)
    idx = idx[idx + step < flat.shape[0]]
    tmp = flat[idx].copy()
    flat[idx] = flat[idx + 1]
    flat[idx + 1] = tmp
    return a


# ---------------------------------------------------------------- bootstrap

def stage_outputs(poly, xext_dev, stream=None):
    """Every intermediate of the plain path for host polynomials (rows, N, 4) and a device
    xext_fft: the launch sequence of fk20_poly2h_fft.cu:57-77, one output kept per stage."""
    poly = np.ascontiguousarray(poly, dtype=np.uint64)
    rows, N = poly.shape[:2]
    n2 = N // 8
    P = L.DeviceBuffer.from_array(poly)
    tc = L.DeviceBuffer(rows * 16 * n2 * 32)
    tcf = L.DeviceBuffer(rows * 16 * n2 * 32)
    hext = L.DeviceBuffer(rows * n2 * 144)
    h = L.DeviceBuffer(rows * n2 * 144)
    hf = L.DeviceBuffer(rows * n2 * 144)
    L.call("fk20_poly2toeplitz_coefficients", tc.ptr, P.ptr, rows, N, stream)
    L.call("fr_fft", tcf.ptr, tc.ptr, rows * 16, n2, stream)
    L.call("fk20_msm", hext.ptr, tcf.ptr, xext_dev.ptr, rows, n2, stream)
    L.call("g1p_ift", h.ptr, hext.ptr, rows, n2, stream)
    L.call("fk20_hext2h", h.ptr, rows, n2, stream)
    L.call("g1p_fft", hf.ptr, h.ptr, rows, n2, stream)
    L.synchronize()
    sh = shapes(rows, N)
    out = {"polynomial": poly}
    for name, buf in (("toeplitz_coefficients", tc), ("toeplitz_coefficients_fft", tcf),
                      ("hext_fft", hext), ("h", h), ("h_fft", hf)):
        out[name] = buf.download(sh[name])
    for b in (P, tc, tcf, hext, h, hf):
        b.free()
    return out


def bootstrap(path, rows=REFERENCE_ROWS, N=REFERENCE_N, seed=512, tau=None, overwrite=False):
    """Write a known-answer file of `rows` seeded random polynomials and a setup of secret tau
    (seeded when None), computed on the GPU.  Returns (bytes written, tau)."""
    if tau is None:
        tau = S.random_secret(seed + 1)
    xext = S.device_setup(N, tau)
    arrays = stage_outputs(S.random_polys(rows, N, seed), xext)
    arrays["xext_fft"] = xext.download((16, N // 8, 18))
    xext.free()
    return write(path, arrays, overwrite=overwrite), tau


# ---------------------------------------------------------------- verification

def _count_mismatches(kind, got, expected, count):
    """Elements of got (a DeviceBuffer) that differ from the host array expected: fr modulo r,
    g1p as projective points."""
    exp = L.DeviceBuffer.from_array(np.ascontiguousarray(expected, dtype=np.uint64))
    try:
        if kind == "fr":
            L.call("fk20_fr_op", 6, got.ptr, got.ptr, None, count, None)  # canonical form
            L.call("fk20_fr_op", 6, exp.ptr, exp.ptr, None, count, None)
            L.synchronize()
            a, b = got.download((count, 4)), exp.download((count, 4))
            return int(np.any(a != b, axis=1).sum())
        eq = L.DeviceBuffer(count * 144)
        try:
            L.call("fk20_g1p_op", 7, eq.ptr, got.ptr, exp.ptr, None, count, None)
            L.synchronize()
            return int(count - eq.download((count, 18))[:, 0].sum())
        finally:
            eq.free()
    finally:
        exp.free()


def _run_stage(name, rows, N, inputs, prover):
    """Run stage `name` on host inputs (dict of arrays); returns (kind, output DeviceBuffer, count)."""
    n2 = N // 8
    dev = {key: L.DeviceBuffer.from_array(np.ascontiguousarray(v, dtype=np.uint64)) for key, v in inputs.items()}
    fr_out = lambda: L.DeviceBuffer(rows * 16 * n2 * 32)  # noqa: E731
    g1_out = lambda: L.DeviceBuffer(rows * n2 * 144)  # noqa: E731
    try:
        if name == "toeplitz_coefficients2toeplitz_coefficients_fft":
            out = fr_out()
            L.call("fr_fft", out.ptr, dev["toeplitz_coefficients"].ptr, rows * 16, n2, None)
            return "fr", out, rows * 16 * n2
        if name == "poly2toeplitz_coefficients":
            out = fr_out()
            L.call("fk20_poly2toeplitz_coefficients", out.ptr, dev["polynomial"].ptr, rows, N, None)
            return "fr", out, rows * 16 * n2
        out = g1_out()
        if name == "h2h_fft":
            L.call("g1p_fft", out.ptr, dev["h"].ptr, rows, n2, None)
        elif name == "h_fft2h":
            L.call("g1p_ift", out.ptr, dev["h_fft"].ptr, rows, n2, None)
        elif name == "hext_fft2h":
            L.call("g1p_ift", out.ptr, dev["hext_fft"].ptr, rows, n2, None)
            L.call("fk20_hext2h", out.ptr, rows, n2, None)
        elif name == "hext_fft2h_fft":
            L.call("fk20_hext_fft2h_fft", out.ptr, dev["hext_fft"].ptr, rows, n2, None)
        elif name == "msm":
            L.call("fk20_msm", out.ptr, dev["toeplitz_coefficients_fft"].ptr, dev["xext_fft"].ptr, rows, n2, None)
        elif name == "poly2hext_fft":
            tmp = fr_out()
            L.call("fk20_poly2toeplitz_coefficients", tmp.ptr, dev["polynomial"].ptr, rows, N, None)
            L.call("fr_fft", tmp.ptr, tmp.ptr, rows * 16, n2, None)
            L.call("fk20_msm", out.ptr, tmp.ptr, dev["xext_fft"].ptr, rows, n2, None)
            L.synchronize()
            tmp.free()
        elif name == "poly2h_fft":
            L.call("fk20_poly2h_fft", out.ptr, dev["polynomial"].ptr, dev["xext_fft"].ptr, rows, N, None)
        elif name == "prove_rows":
            sid = prover.load_setup(dev["xext_fft"], N, 8)
            try:
                prover.prove_device(sid, out, dev["polynomial"], rows)
                L.synchronize()
            finally:
                prover.unload_setup(sid)
        else:
            raise ValueError("unknown stage %r" % name)
        return "g1", out, rows * n2
    finally:
        L.synchronize()
        for b in dev.values():
            b.free()


# (stage, inputs, expected output, the array mangled in the falsifiability run and the
# varMangle step, both as fk20_512test.cu:399-857 chooses them): the stage tests of
# fk20_512test.cu:93-121 in its order; prove_rows is the batched path (fk20_prove_rows) in
# the place of the comb MSM tests (fk20_msmcomb_512).  The NTT test mangles its expected
# output (:399): its input's swapped slots 3 and 4 are zero in every Toeplitz vector.
STAGES = (
    ("toeplitz_coefficients2toeplitz_coefficients_fft", ("toeplitz_coefficients",),
     "toeplitz_coefficients_fft", "toeplitz_coefficients_fft", 512),
    ("h2h_fft", ("h",), "h_fft", "h", 128),
    ("h_fft2h", ("h_fft",), "h", "h_fft", 128),
    ("hext_fft2h", ("hext_fft",), "h", "hext_fft", 128),
    ("hext_fft2h_fft", ("hext_fft",), "h_fft", "hext_fft", 32),
    ("poly2toeplitz_coefficients", ("polynomial",), "toeplitz_coefficients", "polynomial", 8),
    ("poly2hext_fft", ("polynomial", "xext_fft"), "hext_fft", "xext_fft", 32),
    ("msm", ("toeplitz_coefficients_fft", "xext_fft"), "hext_fft", "toeplitz_coefficients_fft", 512),
    ("poly2h_fft", ("polynomial", "xext_fft"), "h_fft", "xext_fft", 32),
    ("prove_rows", ("polynomial", "xext_fft"), "h_fft", "polynomial", 64),
)
STAGE_NAMES = tuple(s[0] for s in STAGES)


def verify(kat, rows=None, N=REFERENCE_N, stages=STAGE_NAMES, falsify=False):
    """Run the stage tests on the first `rows` rows of a known-answer file (a path or the dict
    `read` returns).  falsify: each stage's input (or expected output) is mangled first, and
    the stage passes only if some output differs (fk20_512test.cu:267-366).  Returns one dict per stage with the
    element count, the mismatch count, pass/fail and the stage's wall time."""
    if not isinstance(kat, dict):
        kat = read(kat, N)
    total = kat["polynomial"].shape[0]
    N = kat["polynomial"].shape[1]
    rows = total if rows is None else rows
    if not 1 <= rows <= total:
        raise ValueError("rows must be in [1, %d]" % total)
    prover = None
    if "prove_rows" in stages:
        from .prover import Prover
        prover = Prover()
    report = []
    try:
        for name, ins, exp_name, mangled, step in STAGES:
            if name not in stages:
                continue
            inputs = {}
            for key in ins:
                a = kat[key] if key == "xext_fft" else kat[key][:rows]
                inputs[key] = mangle(a, step) if falsify and key == mangled else np.asarray(a)
            expected = kat[exp_name][:rows]
            if falsify and mangled == exp_name:
                expected = mangle(expected, step)
            t0 = time.perf_counter()
            kind, out, count = _run_stage(name, rows, N, inputs, prover)
            try:
                bad = _count_mismatches(kind, out, expected, count)
            finally:
                out.free()
            ok = bad > 0 if falsify else bad == 0
            report.append({"stage": name, "elements": count, "mismatches": bad, "pass": ok,
                           "s": round(time.perf_counter() - t0, 3)})
    finally:
        if prover is not None:
            prover.close()
    return report
