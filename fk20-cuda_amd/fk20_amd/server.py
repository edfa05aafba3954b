"""Serving loop front end: fk20_server_* (continuous batching over a Prover context)."""
import ctypes

from .lib import call, lib


class Server:
    def __init__(self, prover, budget_bytes=8 << 30, slots=2, max_batch_rows=2048):
        s = ctypes.c_void_p()
        call("fk20_server_create", ctypes.byref(s), prover.ctx, budget_bytes, slots, max_batch_rows)
        self.handle = s.value

    def submit(self, setup_id, poly_ptr, rows, out_ptr):
        rid = ctypes.c_uint64()
        call("fk20_server_submit", self.handle, setup_id, poly_ptr, rows, out_ptr, ctypes.byref(rid))
        return rid.value

    def wait(self, rid):
        lat = ctypes.c_double()
        call("fk20_server_wait", self.handle, rid, ctypes.byref(lat))
        return lat.value

    def stats(self):
        b, r, m = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        status = lib().fk20_server_stats(self.handle, ctypes.byref(b), ctypes.byref(r), ctypes.byref(m))
        return {"batches": b.value, "rows_done": r.value, "max_inflight_rows": m.value, "status": status}

    def close(self):
        if self.handle:
            call("fk20_server_destroy", self.handle)
            self.handle = None
