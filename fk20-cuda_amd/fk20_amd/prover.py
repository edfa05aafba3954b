"""Batched FK20 prover: the context API of libfk20 (fk20_ctx_*, fk20_setup_load,
fk20_prove_rows) with device buffers managed from Python."""
import ctypes

import numpy as np

from .lib import DeviceBuffer, call, lib, synchronize


class BatchItem(ctypes.Structure):
    """fk20_batch_item_t of include/fk20.h."""
    _fields_ = [("setup_id", ctypes.c_int), ("rows", ctypes.c_uint),
                ("polynomial", ctypes.c_void_p), ("h_fft", ctypes.c_void_p)]


MAX_ITEMS = 16


class Prover:
    """One libfk20 context on the current device; setups stay resident until closed."""

    def __init__(self, graph=False):
        c = ctypes.c_void_p()
        call("fk20_ctx_create", ctypes.byref(c))
        self.ctx = c.value
        self.setups = {}
        if graph:
            call("fk20_ctx_set_graph", self.ctx, 1)

    def load_setup(self, xext_fft, N, window=8, stream=None):
        """xext_fft: host array (16, N/8, 18) or a DeviceBuffer holding it."""
        buf = xext_fft if isinstance(xext_fft, DeviceBuffer) else DeviceBuffer.from_array(
            np.ascontiguousarray(xext_fft, dtype=np.uint64))
        sid = ctypes.c_int()
        try:
            call("fk20_setup_load", self.ctx, buf.ptr, N, window, ctypes.byref(sid), stream)
        finally:
            if buf is not xext_fft:
                buf.free()
        self.setups[sid.value] = N
        return sid.value

    def load_commit_setup(self, srs, N, window=8, stream=None):
        """KZG commitment setup from the first N SRS points: host (>= N, 18) or a DeviceBuffer."""
        buf = srs if isinstance(srs, DeviceBuffer) else DeviceBuffer.from_array(
            np.ascontiguousarray(srs[:N], dtype=np.uint64))
        sid = ctypes.c_int()
        try:
            call("fk20_commit_setup_load", self.ctx, buf.ptr, N, window, ctypes.byref(sid), stream)
        finally:
            if buf is not srs:
                buf.free()
        self.setups[sid.value] = N
        return sid.value

    def commit_device(self, sid, out, poly, rows, stream=None):
        """Enqueue commitments of rows: poly (rows, N) fr_t and out (rows) g1p_t device buffers."""
        op = out.ptr if isinstance(out, DeviceBuffer) else out
        pp = poly.ptr if isinstance(poly, DeviceBuffer) else poly
        call("fk20_commit_rows", self.ctx, sid, op, pp, rows, stream)

    def commit(self, sid, poly):
        """Host polynomials (rows, N, 4) -> host commitments (rows, 18)."""
        poly = np.ascontiguousarray(poly, dtype=np.uint64)
        rows, N = poly.shape[0], poly.shape[1]
        assert self.setups[sid] == N
        pb = DeviceBuffer.from_array(poly)
        out = DeviceBuffer(max(1, rows) * 144)
        self.commit_device(sid, out, pb, rows)
        synchronize()
        res = out.download((rows, 18))
        pb.free()
        out.free()
        return res

    def unload_setup(self, sid):
        call("fk20_setup_unload", self.ctx, sid)
        self.setups.pop(sid, None)

    def reserve(self, rows, N):
        call("fk20_ctx_reserve", self.ctx, rows, N)

    def prove_device(self, sid, h_fft, poly, rows, stream=None):
        """Enqueue rows on stream: poly and h_fft are DeviceBuffers (or raw pointers)."""
        hp = h_fft.ptr if isinstance(h_fft, DeviceBuffer) else h_fft
        pp = poly.ptr if isinstance(poly, DeviceBuffer) else poly
        call("fk20_prove_rows", self.ctx, sid, hp, pp, rows, stream)

    def prove_batch_device(self, items, stream=None):
        """items: [(setup_id, rows, poly_ptr_or_buffer, h_fft_ptr_or_buffer)] in one batched pass."""
        arr = (BatchItem * max(1, len(items)))()
        for k, (sid, rows, poly, out) in enumerate(items):
            arr[k].setup_id = sid
            arr[k].rows = rows
            arr[k].polynomial = poly.ptr if isinstance(poly, DeviceBuffer) else poly
            arr[k].h_fft = out.ptr if isinstance(out, DeviceBuffer) else out
        call("fk20_prove_batch", self.ctx, arr, len(items), stream)

    def set_output(self, fmt):
        """Output format of the proofs: one of OUTPUT_FORMATS (default "g1p")."""
        call("fk20_ctx_set_output", self.ctx, OUTPUT_FORMATS[fmt])
        self.out_format = fmt

    def _out_shape(self, rows, N):
        fmt = getattr(self, "out_format", "g1p")
        if fmt.startswith("compressed"):
            return (rows, N // 8, 48), np.uint8
        return (rows, N // 8, 18), np.uint64

    def prove_batch(self, groups):
        """groups: [(setup_id, host polys (rows, N, 4))] -> [host h_fft (rows, N/8, 18)], or
        (rows, N/8, 48) uint8 in the compressed output formats."""
        bufs, items = [], []
        for sid, poly in groups:
            poly = np.ascontiguousarray(poly, dtype=np.uint64)
            rows, N = poly.shape[0], poly.shape[1]
            assert self.setups[sid] == N
            pb = DeviceBuffer.from_array(poly)
            shape, dt = self._out_shape(rows, N)
            hb = DeviceBuffer(int(np.prod(shape)) * np.dtype(dt).itemsize)
            bufs.append((pb, hb, shape, dt))
            items.append((sid, rows, pb, hb))
        self.prove_batch_device(items)
        synchronize()
        outs = []
        for pb, hb, shape, dt in bufs:
            outs.append(hb.download(shape, dt))
            pb.free()
            hb.free()
        return outs

    def prove(self, sid, poly):
        """Host polynomials (rows, N, 4) -> host h_fft (rows, N/8, 18)."""
        poly = np.ascontiguousarray(poly, dtype=np.uint64)
        rows, N = poly.shape[0], poly.shape[1]
        assert self.setups[sid] == N
        pb = DeviceBuffer.from_array(poly)
        shape, dt = self._out_shape(rows, N)
        hb = DeviceBuffer(int(np.prod(shape)) * np.dtype(dt).itemsize)
        self.prove_device(sid, hb, pb, rows)
        synchronize()
        out = hb.download(shape, dt)
        pb.free()
        hb.free()
        return out

    def set_streams(self, n):
        """1: one launch sequence on the caller's stream; 2 (default): the rows split in two
        halves on two internal streams forked from and joined to the caller's stream."""
        call("fk20_ctx_set_streams", self.ctx, int(n))

    def set_schedule(self, wnaf=5, radix=4, waves_per_simd=2):
        """Schedule of the transform kernels: wNAF width 3..5 of the twiddle programs, radix 4
        (radix-2^2 butterflies and twiddle layers) or 2, waves per SIMD 2 or 3."""
        call("fk20_ctx_set_schedule", self.ctx, int(wnaf), int(radix), int(waves_per_simd))

    def schedule(self):
        w, r, o = ctypes.c_uint(), ctypes.c_uint(), ctypes.c_uint()
        call("fk20_ctx_get_schedule", self.ctx, ctypes.byref(w), ctypes.byref(r), ctypes.byref(o))
        return {"wnaf": w.value, "radix": r.value, "waves_per_simd": o.value}

    def set_chain(self, enable):
        """Transform stages as one persistent launch, or one launch per stage (default)."""
        call("fk20_ctx_set_chain", self.ctx, int(bool(enable)))

    def set_chain_wait(self, ticks):
        """Bound of each wait of the persistent transform launch, in wall-clock ticks."""
        call("fk20_ctx_set_chain_wait", self.ctx, int(ticks))

    def set_graph(self, enable):
        call("fk20_ctx_set_graph", self.ctx, int(bool(enable)))

    def set_launch_timing(self, enable):
        call("fk20_ctx_set_launch_timing", self.ctx, int(bool(enable)))

    def launch_times(self, max_launches=64):
        """[(kind, ms)] of the last fk20_prove_rows call made with launch timing on."""
        ms = (ctypes.c_float * max_launches)()
        kinds = (ctypes.c_int * max_launches)()
        n = ctypes.c_int()
        call("fk20_ctx_launch_times", self.ctx, ms, kinds, max_launches, ctypes.byref(n))
        return [(kinds[i], ms[i]) for i in range(n.value)]

    def close(self):
        if self.ctx:
            call("fk20_ctx_destroy", self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def setup_table_bytes(N, window):
    return int(lib().fk20_setup_table_bytes(N, window))


def twiddle_selfcheck():
    """(status, program size in ops, wNAF digit width) of the library's twiddle programs."""
    n, w = ctypes.c_uint(), ctypes.c_uint()
    status = lib().fk20_twiddle_selfcheck(ctypes.byref(n), ctypes.byref(w))
    return status, n.value, w.value


def twiddle_program(which, idx, wnaf=5):
    """The ops (uint32 array) of the program the transform kernels run for w^idx (which 0),
    w^-idx (which 1) or 2^-idx (which 2) at wNAF width wnaf; host-only, no device needed."""
    n = ctypes.c_uint()
    status = lib().fk20_twiddle_program(which, idx, wnaf, None, 0, ctypes.byref(n))
    if status:
        raise ValueError("fk20_twiddle_program(%d, %d): status %d" % (which, idx, status))
    ops = np.zeros(n.value, dtype=np.uint32)
    status = lib().fk20_twiddle_program(which, idx, wnaf, ops.ctypes.data_as(ctypes.c_void_p), n.value,
                                        ctypes.byref(n))
    if status:
        raise ValueError("fk20_twiddle_program(%d, %d): status %d" % (which, idx, status))
    return ops


# fk20_ctx_set_output formats (include/fk20.h FK20_OUT_*)
OUTPUT_FORMATS = {"g1p": 0, "g1p_da": 1, "compressed": 2, "compressed_da": 3}

LAUNCH_KINDS = {0: "k_toeplitz_fft_T", 1: "k_msm_comb", 2: "k_dit_stage", 3: "k_middle", 4: "k_dif_stage",
                5: "k_fft_output", 6: "k_transform_chain", 7: "k_twiddle_layer"}
