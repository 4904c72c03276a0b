"""Loader for libeithot.so (the HIP build for gfx950).

The library is built in-tree by `__graft_entry__.build()` (or `make -C
eithot_amd/csrc`).  There is no CPU fallback: if the library is missing,
every entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libeithot.so")

EITHOT_OK = 0
ERRORS = {
    -1: "EITHOT_ERR_INVALID",
    -2: "EITHOT_ERR_ELEM",
    -3: "EITHOT_ERR_ALPHA",
    -4: "EITHOT_ERR_DEVICE",
    -5: "EITHOT_ERR_NOMEM",
    -6: "EITHOT_ERR_UNSUPPORTED",
    -7: "EITHOT_ERR_NO_DEVICE",
    -8: "EITHOT_ERR_BUSY",
}

# every symbol include/eithot.h declares
EXPORTS = [
    "eithot_transpose", "eithot_transpose_async", "eithot_transpose_noalloc",
    "eithot_inplace_transpose",
    "eithot_last_peak_extra_bytes", "eithot_last_error", "eithot_release", "eithot_estimate",
    "eithot_estimate_reference", "eithot_decompose", "eithot_plan_describe",
    "eithot_plan_partition", "eithot_plan_block_table", "eithot_plan_cycle_maps",
    "eithot_plan_segments",
    "eithot_set_tuning",
    "eithot_get_tuning",
    "eithot_set_pass_timing", "eithot_last_pass_times", "eithot_fill_pattern",
    "eithot_verify_pattern", "eithot_batch_create", "eithot_batch_add", "eithot_batch_build",
    "eithot_batch_run", "eithot_batch_info", "eithot_batch_destroy", "eithot_server_create",
    "eithot_server_submit", "eithot_server_poll", "eithot_server_status", "eithot_server_drain",
    "eithot_server_destroy",
]


class EithotError(RuntimeError):
    def __init__(self, code, msg=""):
        self.code = code
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")


class OpDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("map", ctypes.c_int), ("staged", ctypes.c_int),
                ("stage", ctypes.c_int), ("base", ctypes.c_uint64)] + [
                    (n, ctypes.c_int64) for n in
                    ("B", "R", "C", "U", "us", "W", "m", "n", "c", "a", "b", "k", "binv")]


class RefProfile(ctypes.Structure):
    _fields_ = [("smem_per_block", ctypes.c_int64), ("sms", ctypes.c_int),
                ("max_threads_per_sm", ctypes.c_int), ("padding_blocks_per_sm", ctypes.c_int)]


_LIB = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP extension first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    # One HIP runtime per process: PyTorch ships its own, and a process that
    # loads libeithot.so first and torch afterwards ends up with two, one of
    # which finds no device.  Torch, when installed, is loaded first so the
    # library binds to the runtime torch's tensors come from.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    i64p = ctypes.POINTER(ctypes.c_int64)
    ip = ctypes.POINTER(ctypes.c_int)
    dp = ctypes.POINTER(ctypes.c_double)
    vp = ctypes.c_void_p
    sz = ctypes.c_size_t
    L.eithot_transpose.argtypes = [vp, i64p, ip, ctypes.c_int, sz, ctypes.c_int, ctypes.c_double]
    L.eithot_transpose_async.argtypes = [vp, i64p, ip, ctypes.c_int, sz, ctypes.c_int,
                                         ctypes.c_double, vp]
    L.eithot_transpose_noalloc.argtypes = [vp, i64p, ip, ctypes.c_int, sz, ctypes.c_int,
                                           ctypes.c_double]
    L.eithot_inplace_transpose.argtypes = [vp, ctypes.c_int, ctypes.c_int, ip, ip, sz,
                                           ctypes.c_int, ctypes.c_double]
    L.eithot_last_peak_extra_bytes.restype = sz
    L.eithot_last_error.restype = ctypes.c_char_p
    L.eithot_estimate.argtypes = [i64p, ip, ctypes.c_int, sz, ctypes.c_double, ip, dp,
                                  ctypes.POINTER(sz)]
    L.eithot_estimate_reference.argtypes = [i64p, ip, ctypes.c_int, sz, ctypes.c_double,
                                            ctypes.POINTER(RefProfile), dp, dp, dp, dp, ip, dp,
                                            ip]
    L.eithot_decompose.argtypes = [i64p, ip, ctypes.c_int, ctypes.c_int, ip, ip, i64p,
                                   ctypes.c_int]
    L.eithot_plan_describe.argtypes = [i64p, ip, ctypes.c_int, sz, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_int64, ctypes.POINTER(OpDesc), ctypes.c_int, i64p]
    L.eithot_plan_partition.argtypes = [i64p, ip, ctypes.c_int, sz, ctypes.c_int,
                                        ctypes.c_double, ip, ip, i64p, i64p]
    L.eithot_plan_block_table.restype = ctypes.c_int64
    L.eithot_plan_block_table.argtypes = [i64p, ip, ctypes.c_int, sz, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_int64, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_uint32), ctypes.c_int64]
    L.eithot_plan_cycle_maps.argtypes = [i64p, ip, ctypes.c_int, sz, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_int64, ctypes.c_int, i64p, ctypes.c_int]
    L.eithot_plan_segments.argtypes = [i64p, ip, ctypes.c_int, sz, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_int64, ctypes.c_int, i64p, i64p, i64p]
    L.eithot_set_tuning.argtypes = [ctypes.c_char_p, ctypes.c_int64]
    L.eithot_get_tuning.argtypes = [ctypes.c_char_p, i64p]
    L.eithot_set_pass_timing.argtypes = [ctypes.c_int]
    L.eithot_last_pass_times.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_int]
    L.eithot_fill_pattern.argtypes = [vp, ctypes.c_uint64, sz, ctypes.c_uint64, ctypes.c_int,
                                      ctypes.c_uint64, vp]
    L.eithot_verify_pattern.argtypes = [vp, i64p, ip, ctypes.c_int, sz, ctypes.c_uint64,
                                        ctypes.c_int, ctypes.POINTER(ctypes.c_ulonglong)]
    L.eithot_batch_create.restype = vp
    L.eithot_batch_add.argtypes = [vp, vp, i64p, ip, ctypes.c_int, sz]
    L.eithot_batch_build.argtypes = [vp, ctypes.c_double, ctypes.c_int]
    L.eithot_batch_run.argtypes = [vp, vp, ctypes.c_int]
    L.eithot_batch_info.argtypes = [vp, ip, i64p, ip]
    L.eithot_batch_destroy.argtypes = [vp]
    L.eithot_server_create.restype = vp
    L.eithot_server_create.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int]
    L.eithot_server_submit.restype = ctypes.c_int64
    L.eithot_server_submit.argtypes = [vp, vp, i64p, ip, ctypes.c_int, sz]
    L.eithot_server_poll.argtypes = [vp, ctypes.c_int]
    L.eithot_server_status.argtypes = [vp, ctypes.c_int64, dp]
    L.eithot_server_drain.argtypes = [vp]
    L.eithot_server_destroy.argtypes = [vp]
    _LIB = L
    return L


def check(rc):
    if rc < 0:
        msg = lib().eithot_last_error().decode(errors="replace")
        raise EithotError(rc, msg)
    return rc


def i64(seq):
    return (ctypes.c_int64 * len(seq))(*[int(x) for x in seq])


def i32(seq):
    return (ctypes.c_int * len(seq))(*[int(x) for x in seq])
