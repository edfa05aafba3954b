"""ctypes binding of libfk20.so (the C ABI in include/fk20.h).

The library is loaded from the package directory (fk20-cuda_amd/libfk20.so).  If it is
missing, loading raises: there is no CPU fallback for any product entry point.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# FK20_LIBRARY selects another build of the library (e.g. a variant built for comparison)
LIBPATH = os.environ.get("FK20_LIBRARY", os.path.join(PKG_DIR, "libfk20.so"))

FK20_OK = 0
STATUS = {0: "ok", -1: "invalid argument", -2: "out of device memory", -3: "HIP runtime error",
          -4: "no gfx950 device", -5: "a bounded device-side wait expired",
          -6: "request larger than the fixed workspace of an allocation-free entry point"}


class Fk20Error(RuntimeError):
    def __init__(self, fn, status, hip_error=0):
        super().__init__("%s failed: %s (status %d, hip error %d)" % (fn, STATUS.get(status, "?"), status, hip_error))
        self.status = status
        self.hip_error = hip_error


_lib = None

_P = ctypes.c_void_p
_U = ctypes.c_uint
_SZ = ctypes.c_size_t
_I = ctypes.c_int

_SIGS = {
    "fk20_last_hip_error": ([], _I),
    "fk20_poly2toeplitz_coefficients": ([_P, _P, _U, _U, _P], _I),
    "fk20_poly2toeplitz_coefficients_fft": ([_P, _P, _U, _U, _P], _I),
    "fr_fft": ([_P, _P, _U, _U, _P], _I),
    "fr_ift": ([_P, _P, _U, _U, _P], _I),
    "g1p_fft": ([_P, _P, _U, _U, _P], _I),
    "g1p_ift": ([_P, _P, _U, _U, _P], _I),
    "fk20_msm": ([_P, _P, _P, _U, _U, _P], _I),
    "fk20_hext2h": ([_P, _U, _U, _P], _I),
    "fk20_hext_fft2h_fft": ([_P, _P, _U, _U, _P], _I),
    "fk20_poly2h_fft": ([_P, _P, _P, _U, _U, _P], _I),
    "fk20_poly2h_fft_noalloc": ([_P, _P, _P, _U, _U, _P], _I),
    "fk20_setup2xext_fft": ([_P, _P, _U, _P], _I),
    "fk20_g1p_to_affine": ([_P, _P, _SZ, _P], _I),
    "fk20_twiddle_program": ([_I, _U, _U, _P, _U, ctypes.POINTER(_U)], _I),
    "fk20_fp_mul_probe": ([_I, _P, _P, _P, _U, _U, _P], _I),
    "fk20_ctx_set_schedule": ([_P, _U, _U, _U], _I),
    "fk20_ctx_get_schedule": ([_P, ctypes.POINTER(_U), ctypes.POINTER(_U), ctypes.POINTER(_U)], _I),
    "fk20_fp_op": ([_I, _P, _P, _P, _SZ, _P], _I),
    "fk20_fr_op": ([_I, _P, _P, _P, _SZ, _P], _I),
    "fk20_g1p_op": ([_I, _P, _P, _P, _P, _SZ, _P], _I),
    "fk20_ctx_create": ([ctypes.POINTER(_P)], _I),
    "fk20_ctx_destroy": ([_P], _I),
    "fk20_ctx_set_graph": ([_P, _I], _I),
    "fk20_ctx_set_chain": ([_P, _I], _I),
    "fk20_ctx_set_chain_wait": ([_P, ctypes.c_uint64], _I),
    "fk20_ctx_set_streams": ([_P, _U], _I),
    "fk20_ctx_set_output": ([_P, _I], _I),
    "fk20_ctx_get_output": ([_P, ctypes.POINTER(_I)], _I),
    "fk20_output_point_bytes": ([_I], _SZ),
    "fk20_g1p_compress": ([_P, _P, _SZ, _P], _I),
    "fk20_commit_setup_load": ([_P, _P, _U, _U, ctypes.POINTER(_I), _P], _I),
    "fk20_commit_setup_table_bytes": ([_U, _U], _SZ),
    "fk20_commit_workspace_bytes": ([_U, _U], _SZ),
    "fk20_commit_rows": ([_P, _I, _P, _P, _U, _P], _I),
    "fk20_setup_table_bytes": ([_U, _U], _SZ),
    "fk20_setup_load": ([_P, _P, _U, _U, ctypes.POINTER(_I), _P], _I),
    "fk20_setup_unload": ([_P, _I], _I),
    "fk20_rows_workspace_bytes": ([_U, _U], _SZ),
    "fk20_ctx_reserve": ([_P, _U, _U], _I),
    "fk20_prove_rows": ([_P, _I, _P, _P, _U, _P], _I),
    "fk20_prove_batch": ([_P, _P, _U, _P], _I),
    "fk20_prove_batch_ws": ([_P, _P, _U, _P, _SZ, _P], _I),
    "fk20_setup_info": ([_P, _I, ctypes.POINTER(_U), ctypes.POINTER(_U)], _I),
    "fk20_server_create": ([ctypes.POINTER(_P), _P, _SZ, _U, _U], _I),
    "fk20_server_submit": ([_P, _I, _P, _U, _P, ctypes.POINTER(ctypes.c_uint64)], _I),
    "fk20_server_wait": ([_P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)], _I),
    "fk20_server_stats": ([_P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                           ctypes.POINTER(ctypes.c_uint64)], _I),
    "fk20_server_destroy": ([_P], _I),
    "fk20_host_alloc": ([ctypes.POINTER(_P), _SZ], _I),
    "fk20_host_free": ([_P], _I),
    "fk20_twiddle_selfcheck": ([ctypes.POINTER(_U), ctypes.POINTER(_U)], _I),
    "fk20_mad_probe": ([_P, _U, _U, _P], _I),
    "fk20_ctx_set_launch_timing": ([_P, _I], _I),
    "fk20_ctx_launch_times": ([_P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(_I), _I, ctypes.POINTER(_I)], _I),
    "fk20_status_string": ([_I], ctypes.c_char_p),
    "fk20_device_count": ([ctypes.POINTER(_I)], _I),
    "fk20_set_device": ([_I], _I),
    "fk20_malloc": ([ctypes.POINTER(_P), _SZ], _I),
    "fk20_free": ([_P], _I),
    "fk20_memcpy_h2d": ([_P, _P, _SZ], _I),
    "fk20_memcpy_d2h": ([_P, _P, _SZ], _I),
    "fk20_memset": ([_P, _I, _SZ], _I),
    "fk20_memcpy_d2d": ([_P, _P, _SZ], _I),
    "fk20_synchronize": ([], _I),
    "fk20_stream_create": ([ctypes.POINTER(_P)], _I),
    "fk20_stream_destroy": ([_P], _I),
    "fk20_stream_synchronize": ([_P], _I),
    "fk20_event_create": ([ctypes.POINTER(_P)], _I),
    "fk20_event_destroy": ([_P], _I),
    "fk20_event_record": ([_P, _P], _I),
    "fk20_event_elapsed": ([ctypes.POINTER(ctypes.c_float), _P, _P], _I),
}

# entry points declared in include/fk20.h (checked by the CPU test suite)
EXPORTS = sorted(_SIGS)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("libfk20.so not built at %s (run __graft_entry__.build())" % LIBPATH)
        L = ctypes.CDLL(LIBPATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(fn_name, status):
    if status != FK20_OK:
        raise Fk20Error(fn_name, status, lib().fk20_last_hip_error())


def call(fn_name, *args):
    check(fn_name, getattr(lib(), fn_name)(*args))


# ---------------------------------------------------------------- device buffers

class DeviceBuffer:
    """A device allocation made through the library (hipMalloc)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        call("fk20_malloc", ctypes.byref(p), max(self.nbytes, 1))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        call("fk20_memcpy_h2d", self.ptr, a.ctypes.data, a.nbytes)

    def download(self, shape, dtype=np.uint64):
        out = np.empty(shape, dtype=dtype)
        call("fk20_memcpy_d2h", out.ctypes.data, self.ptr, out.nbytes)
        return out

    def fill(self, byte):
        call("fk20_memset", self.ptr, int(byte), self.nbytes)

    def free(self):
        if self.ptr:
            call("fk20_free", self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        call("fk20_stream_create", ctypes.byref(s))
        self.handle = s.value

    def synchronize(self):
        call("fk20_stream_synchronize", self.handle)

    def destroy(self):
        if self.handle:
            call("fk20_stream_destroy", self.handle)
            self.handle = None


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        call("fk20_event_create", ctypes.byref(e))
        self.handle = e.value

    def record(self, stream=None):
        call("fk20_event_record", self.handle, stream.handle if stream else None)

    def elapsed_ms(self, stop):
        ms = ctypes.c_float()
        call("fk20_event_elapsed", ctypes.byref(ms), self.handle, stop.handle)
        return ms.value

    def destroy(self):
        if self.handle:
            call("fk20_event_destroy", self.handle)
            self.handle = None


def device_count():
    n = ctypes.c_int()
    status = lib().fk20_device_count(ctypes.byref(n))
    return n.value if status == FK20_OK else 0


def synchronize():
    call("fk20_synchronize")


class PinnedArray:
    """A numpy view of pinned host memory (hipHostMalloc through the library)."""

    def __init__(self, shape, dtype=np.uint64):
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        call("fk20_host_alloc", ctypes.byref(p), max(nbytes, 1))
        self.ptr = p.value
        buf = (ctypes.c_char * max(nbytes, 1)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=self.dtype, count=int(np.prod(self.shape))).reshape(self.shape)

    def free(self):
        if self.ptr:
            self.array = None
            call("fk20_host_free", self.ptr)
            self.ptr = None
