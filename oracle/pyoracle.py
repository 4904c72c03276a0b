"""ctypes bindings for the oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  The product (eithot_amd) never does.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None


class Stage(ctypes.Structure):
    _fields_ = [("rank", ctypes.c_int), ("perm", ctypes.c_int * 4),
                ("dims", ctypes.c_int64 * 4)]


def build():
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])
    if os.path.isdir("/root/reference/ITNT/src"):
        subprocess.check_call(["make", "-s", "-C", _HERE, "ref"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        i64p = ctypes.POINTER(ctypes.c_int64)
        ip = ctypes.POINTER(ctypes.c_int)
        L.or_permute.argtypes = [ctypes.c_void_p, ctypes.c_void_p, i64p, ip, ctypes.c_int,
                                 ctypes.c_int]
        L.or_permute_mt.argtypes = [ctypes.c_void_p, ctypes.c_void_p, i64p, ip, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int]
        L.or_dest_index.argtypes = [ctypes.c_int64, i64p, ip, ctypes.c_int]
        L.or_dest_index.restype = ctypes.c_int64
        L.or_value.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int]
        L.or_value.restype = ctypes.c_uint64
        L.or_fill.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_uint64,
                              ctypes.c_int, ctypes.c_int64]
        L.or_extended_gcd.argtypes = [ctypes.c_int64, ctypes.c_int64, i64p, i64p]
        L.or_decompose.argtypes = [i64p, ip, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(Stage), ctypes.c_int]
        for name, n in [("or_fn_c2r_rotate", 4), ("or_fn_c2r_row_shuffle", 6),
                        ("or_fn_c2r_col_shuffle", 5), ("or_fn_r2c_col_shuffle", 6),
                        ("or_fn_r2c_rotate", 4)]:
            f = getattr(L, name)
            f.argtypes = [ctypes.c_int64] * n
            f.restype = ctypes.c_int64
        L.or_simulate_reference.argtypes = [ctypes.c_void_p, ctypes.c_int64, i64p, ip,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_double, ctypes.c_int]
        L.or_transpose2d_catanzaro.argtypes = [ctypes.c_void_p] + [ctypes.c_int64] * 4 + [
            ctypes.c_int]
        L.or_transpose2d_direct.argtypes = [ctypes.c_void_p] + [ctypes.c_int64] * 4
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's own gcd/reduced-math sources, or None when not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libref.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        ip = ctypes.POINTER(ctypes.c_int)
        R.ref_extended_gcd.argtypes = [ctypes.c_int, ctypes.c_int, ip, ip]
        R.ref_div32.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.POINTER(ctypes.c_uint)]
        R.ref_div64.argtypes = [ctypes.c_ulonglong, ctypes.c_ulonglong,
                                ctypes.POINTER(ctypes.c_ulonglong)]
        _REF = R
    return _REF


def _i64(seq):
    a = (ctypes.c_int64 * len(seq))(*[int(x) for x in seq])
    return a


def _int(seq):
    return (ctypes.c_int * len(seq))(*[int(x) for x in seq])


def _ptr(arr):
    assert arr.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(arr.ctypes.data)


def permute(arr, dims, perm0, elem=None, nthreads=1):
    """Out-of-place permutation of a flat byte-compatible numpy array."""
    elem = elem or arr.dtype.itemsize
    out = np.empty_like(arr)
    fn = lib().or_permute if nthreads == 1 else None
    if nthreads == 1:
        rc = fn(_ptr(arr), _ptr(out), _i64(dims), _int(perm0), len(dims), elem)
    else:
        rc = lib().or_permute_mt(_ptr(arr), _ptr(out), _i64(dims), _int(perm0), len(dims), elem,
                                 nthreads)
    assert rc == 0, rc
    return out


def dest_index(src, dims, perm0):
    return lib().or_dest_index(int(src), _i64(dims), _int(perm0), len(dims))


def fill(n, elem, seed, mode, first=0):
    """Generator values as a flat uint8 array of n*elem bytes."""
    buf = np.empty(n * elem, dtype=np.uint8)
    lib().or_fill(_ptr(buf), n, elem, seed, mode, first)
    return buf


def value(seed, i, mode):
    return lib().or_value(seed, i, mode)


def extended_gcd(a, b):
    g = ctypes.c_int64()
    m = ctypes.c_int64()
    lib().or_extended_gcd(a, b, ctypes.byref(g), ctypes.byref(m))
    return g.value, m.value


def decompose(dims, perm1, reorder=0):
    out = (Stage * 64)()
    n = lib().or_decompose(_i64(dims), _int(perm1), len(dims), reorder, out, 64)
    assert n >= 0, n
    return [(tuple(out[s].perm[:out[s].rank]), tuple(out[s].dims[:out[s].rank]))
            for s in range(n)]


def simulate_reference(buf, dims, perm1, elem, num_block=1, alpha=0.05, reorder=0,
                       capacity_elems=None):
    cap = capacity_elems if capacity_elems is not None else buf.nbytes // elem
    rc = lib().or_simulate_reference(_ptr(buf), cap, _i64(dims), _int(perm1), len(dims), elem,
                                     num_block, alpha, reorder)
    return rc


def transpose2d_catanzaro(buf, B, R, C, U, method):
    return lib().or_transpose2d_catanzaro(_ptr(buf), B, R, C, U, method)


def transpose2d_direct(buf, B, R, C, U):
    return lib().or_transpose2d_direct(_ptr(buf), B, R, C, U)


fn_c2r_rotate = lambda *a: lib().or_fn_c2r_rotate(*a)  # noqa: E731
fn_c2r_row_shuffle = lambda *a: lib().or_fn_c2r_row_shuffle(*a)  # noqa: E731
fn_c2r_col_shuffle = lambda *a: lib().or_fn_c2r_col_shuffle(*a)  # noqa: E731
fn_r2c_col_shuffle = lambda *a: lib().or_fn_r2c_col_shuffle(*a)  # noqa: E731
fn_r2c_rotate = lambda *a: lib().or_fn_r2c_rotate(*a)  # noqa: E731
