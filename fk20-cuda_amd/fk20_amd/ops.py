"""Host-array front end of libfk20: each function uploads numpy arrays (reference ABI
layout, uint64 limbs), runs the C-ABI entry point of the same name on the GPU and
downloads the result.  Shapes:
  fr arrays  (..., 4)   g1p arrays (..., 18)   g1a arrays (..., 12)
These mirror the reference interface (FK20-CUDA fk20.cuh:20-38, fr.cuh:62-69,
g1.cuh:83-100) with explicit row counts and sizes.
"""
import numpy as np

from .lib import DeviceBuffer, call, synchronize


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def _run(fn, out_shape, inputs, extra, out_dtype=np.uint64, inplace=False):
    bufs = [DeviceBuffer.from_array(_u64(a)) for a in inputs]
    out = bufs[0] if inplace else DeviceBuffer(int(np.prod(out_shape)) * 8)
    call(fn, out.ptr, *[b.ptr for b in bufs], *extra, None)
    synchronize()
    res = out.download(out_shape, out_dtype)
    for b in bufs:
        b.free()
    if not inplace:
        out.free()
    return res


def poly2toeplitz_coefficients(poly):
    poly = _u64(poly)
    rows, N = poly.shape[0], poly.shape[1]
    return _run("fk20_poly2toeplitz_coefficients", (rows, 16, N // 8, 4), [poly], [rows, N])


def poly2toeplitz_coefficients_fft(poly):
    poly = _u64(poly)
    rows, N = poly.shape[0], poly.shape[1]
    return _run("fk20_poly2toeplitz_coefficients_fft", (rows, 16, N // 8, 4), [poly], [rows, N])


def fr_fft(x, inverse=False, inplace=False):
    x = _u64(x)
    count, n = x.shape[0], x.shape[1]
    return _run("fr_ift" if inverse else "fr_fft", x.shape, [x], [count, n], inplace=inplace)


def g1p_fft(x, inverse=False, inplace=False):
    x = _u64(x)
    count, n = x.shape[0], x.shape[1]
    return _run("g1p_ift" if inverse else "g1p_fft", x.shape, [x], [count, n], inplace=inplace)


def msm(tc_fft, xext_fft):
    tc_fft = _u64(tc_fft)
    rows, n2 = tc_fft.shape[0], tc_fft.shape[2]
    return _run("fk20_msm", (rows, n2, 18), [tc_fft, xext_fft], [rows, n2])


def hext2h(h):
    h = _u64(h)
    buf = DeviceBuffer.from_array(h)
    call("fk20_hext2h", buf.ptr, h.shape[0], h.shape[1], None)
    synchronize()
    return buf.download(h.shape)


def hext_fft2h_fft(hext_fft):
    hext_fft = _u64(hext_fft)
    rows, n2 = hext_fft.shape[0], hext_fft.shape[1]
    return _run("fk20_hext_fft2h_fft", hext_fft.shape, [hext_fft], [rows, n2])


def poly2h_fft(poly, xext_fft):
    poly = _u64(poly)
    rows, N = poly.shape[0], poly.shape[1]
    return _run("fk20_poly2h_fft", (rows, N // 8, 18), [poly, xext_fft], [rows, N])


def setup2xext_fft(srs, N):
    return _run("fk20_setup2xext_fft", (16, N // 8, 18), [srs], [N])


def to_affine(pts):
    pts = _u64(pts)
    count = pts.size // 18
    return _run("fk20_g1p_to_affine", pts.shape[:-1] + (12,), [pts], [count])


def compress(pts):
    """48-byte compressed serialisation of g1p points (..., 18) -> uint8 (..., 48)."""
    pts = _u64(pts)
    count = pts.size // 18
    buf = DeviceBuffer.from_array(pts)
    out = DeviceBuffer(max(1, count) * 48)
    call("fk20_g1p_compress", out.ptr, buf.ptr, count, None)
    synchronize()
    res = out.download(pts.shape[:-1] + (48,), np.uint8)
    buf.free()
    out.free()
    return res


FP_OPS = {"add": 0, "sub": 1, "mul": 2, "sqr": 3, "neg": 4, "inv": 5, "reduce": 6}
FR_OPS = FP_OPS
G1_OPS = {"add": 0, "sub": 1, "dbl": 2, "mul": 3, "neg": 4, "add_affine": 5, "is_point": 6, "eq": 7}


def _op(fn, code, width, a, b):
    a = _u64(a)
    count = a.size // width
    bufs = [DeviceBuffer.from_array(a)]
    bptr = None
    if b is not None:
        bufs.append(DeviceBuffer.from_array(_u64(b)))
        bptr = bufs[1].ptr
    out = DeviceBuffer(count * width * 8)
    call(fn, code, out.ptr, bufs[0].ptr, bptr, count, None)
    synchronize()
    res = out.download(a.shape)
    for x in bufs + [out]:
        x.free()
    return res


def fp_op(op, a, b=None):
    return _op("fk20_fp_op", FP_OPS[op], 6, a, b)


def fr_op(op, a, b=None):
    return _op("fk20_fr_op", FR_OPS[op], 4, a, b)


def g1p_op(op, p, q=None, k=None):
    p = _u64(p)
    count = p.size // 18
    bufs = [DeviceBuffer.from_array(p)]
    qp = kp = None
    if q is not None:
        bufs.append(DeviceBuffer.from_array(_u64(q)))
        qp = bufs[-1].ptr
    if k is not None:
        bufs.append(DeviceBuffer.from_array(_u64(k)))
        kp = bufs[-1].ptr
    out = DeviceBuffer(count * 18 * 8)
    call("fk20_g1p_op", G1_OPS[op], out.ptr, bufs[0].ptr, qp, kp, count, None)
    synchronize()
    res = out.download(p.shape)
    for x in bufs + [out]:
        x.free()
    return res
