"""eithot_amd -- in-place high-order tensor transposition on MI355X (gfx950).

Host-side mirror of the reference's operator interface over the C-ABI of
libeithot.so:

* ``transpose(data, source, rank, dim, permutation, sizeofType, num_block,
  LARGE_RATIO)`` -- the reference library call ``inplace::transpose``
  (ITNT/include/transpose.h:8): 0-based permutation, element offset
  ``source``, ``dim[0]`` fastest.
* ``transpose_tensor(t, dims, perm1, ...)`` -- the main C-ABI entry point
  ``eithot_transpose`` on a device tensor, 1-based permutation.
* ``estimate(...)`` / ``estimate_reference(...)`` -- the Nd_estimation planner.
* ``Batch`` and ``Server`` -- batched and served transposition.

Device memory and streams come from PyTorch; the transposition itself always
runs the HIP kernels of libeithot.so (there is no CPU path).
"""
import ctypes

from . import _lib
from ._lib import EithotError, check, i32, i64, lib

__all__ = ["transpose", "transpose_tensor", "transpose_ptr", "estimate", "estimate_reference",
           "decompose", "plan_describe", "plan_partition", "last_peak_extra_bytes", "fill_pattern",
           "verify_pattern", "set_tuning", "get_tuning", "Batch", "Server", "EithotError"]


def _ptr_of(data):
    """Device address of a torch tensor or an integer pointer."""
    if hasattr(data, "data_ptr"):
        if not data.is_cuda:
            raise EithotError(-1, "tensor must live on the GPU")
        return data.data_ptr()
    return int(data)


def transpose(data, source, rank, dim, permutation, sizeofType, num_block=1, LARGE_RATIO=0.05):
    """Reference-compatible call: ``inplace::transpose`` (ITNT/src/transpose.cpp:21-30).

    ``permutation`` is 0-based; ``source`` is an element offset into ``data``.
    """
    return check(lib().eithot_inplace_transpose(
        ctypes.c_void_p(_ptr_of(data)), int(source), int(rank), i32(dim[:rank]),
        i32(permutation[:rank]), int(sizeofType), int(num_block), float(LARGE_RATIO)))


def transpose_ptr(ptr, dims, perm1, elem_bytes, n_sub=1, alpha=0.05, stream=None):
    """``eithot_transpose`` on a raw device pointer (1-based perm)."""
    if stream is None:
        return check(lib().eithot_transpose(ctypes.c_void_p(int(ptr)), i64(dims), i32(perm1),
                                            len(dims), int(elem_bytes), int(n_sub),
                                            float(alpha)))
    return check(lib().eithot_transpose_async(ctypes.c_void_p(int(ptr)), i64(dims), i32(perm1),
                                              len(dims), int(elem_bytes), int(n_sub),
                                              float(alpha), ctypes.c_void_p(int(stream))))


def transpose_tensor(t, dims, perm1, elem_bytes=None, n_sub=1, alpha=0.05, stream=None):
    """Transposes the device tensor ``t`` (viewed as dims, dims[0] fastest) in place."""
    elem = elem_bytes or t.element_size()
    nbytes = 1
    for d in dims:
        nbytes *= int(d)
    if nbytes * elem > t.numel() * t.element_size():
        raise EithotError(-1, "tensor smaller than dims * elem_bytes")
    return transpose_ptr(_ptr_of(t), dims, perm1, elem, n_sub, alpha, stream)


def last_peak_extra_bytes():
    return int(lib().eithot_last_peak_extra_bytes())


def estimate(dims, perm1, elem_bytes, alpha):
    """Planner for this build's kernels: (n_sub, alpha_needed, scratch_bytes)."""
    n = ctypes.c_int()
    a = ctypes.c_double()
    x = ctypes.c_size_t()
    check(lib().eithot_estimate(i64(dims), i32(perm1), len(dims), int(elem_bytes), float(alpha),
                                ctypes.byref(n), ctypes.byref(a), ctypes.byref(x)))
    return n.value, a.value, x.value


def estimate_reference(dims, perm1, elem_bytes, alpha, smem_per_block=49152, sms=82,
                       max_threads_per_sm=1536, padding_blocks_per_sm=1):
    """The reference Nd_estimation policy on a CUDA device profile (RTX 3090 default)."""
    prof = _lib.RefProfile(smem_per_block, sms, max_threads_per_sm, padding_blocks_per_sm)
    vals = [ctypes.c_double() for _ in range(4)]
    nb = ctypes.c_int()
    a = ctypes.c_double()
    flags = ctypes.c_int()
    check(lib().eithot_estimate_reference(
        i64(dims), i32(perm1), len(dims), int(elem_bytes), float(alpha), ctypes.byref(prof),
        *[ctypes.byref(v) for v in vals], ctypes.byref(nb), ctypes.byref(a), ctypes.byref(flags)))
    return {"init_bytes": vals[0].value, "new_ratio": vals[1].value,
            "final_bytes": vals[2].value, "final_ratio": vals[3].value, "nb": nb.value,
            "alpha": a.value, "printed_new_ratio": bool(flags.value & 1),
            "decomposed": bool(flags.value & 2)}


def decompose(dims, perm1, reorder=0):
    """Decomposer stages as [(perm, dims)] (1-based low-order perms)."""
    n = len(dims)
    rk = (ctypes.c_int * 64)()
    pm = (ctypes.c_int * 256)()
    dm = (ctypes.c_int64 * 256)()
    cnt = check(lib().eithot_decompose(i64(dims), i32(perm1), n, int(reorder), rk, pm, dm, 64))
    return [(tuple(pm[4 * s:4 * s + rk[s]]), tuple(dm[4 * s:4 * s + rk[s]])) for s in range(cnt)]


def plan_describe(dims, perm1, elem_bytes, n_sub=1, alpha=0.05, lds_max=0):
    """Device passes of the plan as dicts, plus its scratch bytes."""
    ops = (_lib.OpDesc * 512)()
    scratch = ctypes.c_int64()
    n = check(lib().eithot_plan_describe(i64(dims), i32(perm1), len(dims), int(elem_bytes),
                                         int(n_sub), float(alpha), int(lds_max), ops, 512,
                                         ctypes.byref(scratch)))
    out = []
    for i in range(min(n, 512)):
        o = ops[i]
        d = {f: getattr(o, f) for f, _ in _lib.OpDesc._fields_}
        if d["kind"] == 5:  # fused block permutation: attach its gather table
            import numpy as np
            tab = np.zeros(d["R"], dtype=np.uint32)
            check(lib().eithot_plan_block_table(
                i64(dims), i32(perm1), len(dims), int(elem_bytes), int(n_sub), float(alpha),
                int(lds_max), i, tab.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), tab.size))
            d["table"] = tab
        if d["kind"] == 4:  # cycle following: attach its cell maps (bc, R, C, m)
            mp = (ctypes.c_int64 * 16)()
            k = check(lib().eithot_plan_cycle_maps(
                i64(dims), i32(perm1), len(dims), int(elem_bytes), int(n_sub), float(alpha),
                int(lds_max), i, mp, 4))
            d["maps"] = [tuple(mp[4 * q:4 * q + 4]) for q in range(k)]
        if d["kind"] == 6:  # segmented tile: attach its segment size and digits
            sb = ctypes.c_int64()
            inner = (ctypes.c_int64 * 12)()
            outer = (ctypes.c_int64 * 12)()
            k = check(lib().eithot_plan_segments(
                i64(dims), i32(perm1), len(dims), int(elem_bytes), int(n_sub), float(alpha),
                int(lds_max), i, ctypes.byref(sb), inner, outer))
            d["segb"] = sb.value
            d["inner"] = [(inner[2 * q], inner[2 * q + 1]) for q in range(k % 16)]
            d["outer"] = [(outer[2 * q], outer[2 * q + 1]) for q in range(k // 16)]
        out.append(d)
    return out, scratch.value


def plan_partition(dims, perm1, elem_bytes, n_sub, alpha=0.05):
    """Sub-tensor partition of the plan: (count, 0-based dim, part_len, last_len)."""
    n = ctypes.c_int()
    d = ctypes.c_int()
    q = ctypes.c_int64()
    r = ctypes.c_int64()
    check(lib().eithot_plan_partition(i64(dims), i32(perm1), len(dims), int(elem_bytes),
                                      int(n_sub), float(alpha), ctypes.byref(n), ctypes.byref(d),
                                      ctypes.byref(q), ctypes.byref(r)))
    return n.value, d.value, q.value, r.value


def set_tuning(key, value):
    """Process-wide planner/launch knob (``eithot_set_tuning``); returns the old value."""
    old = get_tuning(key)
    check(lib().eithot_set_tuning(key.encode(), int(value)))
    return old


def get_tuning(key):
    v = ctypes.c_int64()
    check(lib().eithot_get_tuning(key.encode(), ctypes.byref(v)))
    return v.value


def fill_pattern(t, n, elem_bytes, seed, mode, first=0, stream=None):
    return check(lib().eithot_fill_pattern(ctypes.c_void_p(_ptr_of(t)), int(n), int(elem_bytes),
                                           int(seed), int(mode), int(first),
                                           ctypes.c_void_p(int(stream or 0))))


def verify_pattern(t, dims, perm1, elem_bytes, seed, mode):
    """Number of output elements that differ from the permuted pattern."""
    mm = ctypes.c_ulonglong()
    check(lib().eithot_verify_pattern(ctypes.c_void_p(_ptr_of(t)), i64(dims), i32(perm1),
                                      len(dims), int(elem_bytes), int(seed), int(mode),
                                      ctypes.byref(mm)))
    return mm.value


from .batch import Batch, Server  # noqa: E402
