"""Batched and served transposition over the C-ABI (eithot_batch_*, eithot_server_*)."""
import ctypes

from ._lib import EithotError, check, i32, i64, lib


class Batch:
    """Many independent tensors transposed together: one persistent launch per
    decomposition stage, replayed from a HIP graph."""

    def __init__(self):
        self._h = lib().eithot_batch_create()
        if not self._h:
            raise EithotError(-6, "batch creation failed")
        self.count = 0

    def add(self, ptr, dims, perm1, elem_bytes):
        idx = check(lib().eithot_batch_add(self._h, ctypes.c_void_p(int(ptr)), i64(dims),
                                           i32(perm1), len(dims), int(elem_bytes)))
        self.count += 1
        return idx

    def build(self, alpha=0.05, tune=True):
        return check(lib().eithot_batch_build(self._h, float(alpha), 1 if tune else 0))

    def run(self, stream=None, sync=True):
        return check(lib().eithot_batch_run(self._h, ctypes.c_void_p(int(stream or 0)),
                                            1 if sync else 0))

    def info(self):
        n = ctypes.c_int()
        s = ctypes.c_int64()
        g = ctypes.c_int()
        check(lib().eithot_batch_info(self._h, ctypes.byref(n), ctypes.byref(s), ctypes.byref(g)))
        return {"launches": n.value, "scratch_bytes": s.value, "graph": bool(g.value)}

    def close(self):
        if self._h:
            lib().eithot_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop over device-resident tensors."""

    def __init__(self, scratch_pool_bytes, max_inflight=64, nstreams=4):
        self._h = lib().eithot_server_create(int(scratch_pool_bytes), int(max_inflight),
                                             int(nstreams))
        if not self._h:
            raise EithotError(-6, "server creation failed")

    def submit(self, ptr, dims, perm1, elem_bytes):
        return check(lib().eithot_server_submit(self._h, ctypes.c_void_p(int(ptr)), i64(dims),
                                                i32(perm1), len(dims), int(elem_bytes)))

    def poll(self, wait=False):
        return check(lib().eithot_server_poll(self._h, 1 if wait else 0))

    def status(self, rid):
        lat = ctypes.c_double()
        st = lib().eithot_server_status(self._h, int(rid), ctypes.byref(lat))
        return st, lat.value

    def drain(self):
        return check(lib().eithot_server_drain(self._h))

    def close(self):
        if self._h:
            lib().eithot_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
