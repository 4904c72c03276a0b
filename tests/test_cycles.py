"""The cycle-following pass (OP_CYCLE): the planner's cycle tables, and a
host emulation of the device kernel's step order, against the oracle's
direct transpose."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "_build", "libcycles_host.so")


@pytest.fixture(scope="module")
def ch():
    src = os.path.join(HERE, "cycles_harness.cpp")
    deps = [src] + [os.path.join(HERE, "..", "eithot_amd", "csrc", f) for f in ("plan.cpp", "plan.hpp")]
    if not os.path.exists(SO) or os.path.getmtime(SO) < max(os.path.getmtime(d) for d in deps):
        os.makedirs(os.path.dirname(SO), exist_ok=True)
        # parallel test workers may build at once: each writes its own file
        # and renames it into place atomically
        tmp = f"{SO}.{os.getpid()}.tmp"
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-pthread", "-o", tmp,
                               src, deps[1]])
        os.replace(tmp, SO)
    L = ctypes.CDLL(SO)
    L.ch_run.restype = ctypes.c_int64
    L.ch_run.argtypes = [ctypes.c_int64] * 6 + [ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p,
                                                ctypes.POINTER(ctypes.c_int64)]
    L.ch_check.argtypes = [ctypes.c_int64] * 3
    L.ch_run_maps.restype = ctypes.c_int64
    L.ch_run_maps.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int64)] + [ctypes.c_int64] * 4 + \
        [ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p]
    return L


SHAPES = [(2, 3), (3, 2), (8, 36), (36, 8), (48, 320), (64, 100), (7, 11), (16, 16), (1, 9),
          (1440, 8), (11520, 64),
          # at least 2^20 cells: walked on several threads from hashed boundaries
          (1024, 1536), (4097, 513)]


@pytest.mark.parametrize("R,C", SHAPES)
@pytest.mark.parametrize("lmax", [1 << 30, 64, 5, 1])
def test_cycle_tables_cover_every_moved_position_once(ch, R, C, lmax):
    assert ch.ch_check(R, C, lmax) == 0


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("B,R,C,U,lmax,chunk", [
    (3, 8, 36, 48, 1 << 30, 16), (2, 36, 8, 40, 7, 16), (1, 48, 320, 32, 64, 16),
    (2, 13, 17, 24, 3, 8), (1, 64, 100, 16, 9, 16), (4, 5, 7, 64, 1, 64), (1, 120, 64, 16, 100, 16),
    (1, 1024, 1100, 4, 64, 4), (1, 2049, 1023, 2, 256, 2)])
def test_cycle_emulation_equals_transpose(ch, mode, B, R, C, U, lmax, chunk):
    vol = B * R * C * U
    a = O.fill(vol, 1, 17 + mode, 1)
    want = a.copy()
    assert O.transpose2d_direct(want, B, R, C, U) == 0
    b = a.copy()
    ns = ctypes.c_int64()
    segs = ch.ch_run(B, R, C, U, lmax, chunk, mode, 5, b.ctypes.data, ctypes.byref(ns))
    assert segs >= 0
    assert np.array_equal(b, want), (B, R, C, U, lmax, mode, segs, ns.value)


@pytest.mark.parametrize("mode", [0, 3])
@pytest.mark.parametrize("maps,U,lmax", [
    ([(96, 8, 12, 1), (96, 4, 3, 8)], 16, 7),             # two transposes of one 96-cell block
    ([(30, 5, 6, 1), (60, 2, 3, 10), (120, 4, 30, 1)], 8, 1 << 30),
    ([(64, 64, 1, 1), (64, 8, 8, 1)], 32, 5),             # identity map, then a square one
    ([(240, 16, 15, 1), (240, 15, 16, 1)], 16, 11),       # a transpose and its inverse
    ([(1048576, 1024, 1024, 1), (1048576, 32, 2048, 16)], 4, 256)])  # 2^21 cells, threaded walk
def test_composite_cycle_maps_equal_sequential_transposes(ch, mode, maps, U, lmax):
    N = max(bc for bc, _, _, _ in maps) * 2
    vol = N * U
    a = O.fill(vol, 1, 23, 1)
    want = a.copy()
    for bc, R, C, m in maps:  # the passes in order, each a batched transpose
        assert O.transpose2d_direct(want, N // bc, R, C, m * U) == 0
    flat = (ctypes.c_int64 * (4 * len(maps)))(*[v for mp in maps for v in mp])
    b = a.copy()
    assert ch.ch_run_maps(len(maps), flat, N, U, lmax, U, mode, 3, b.ctypes.data) >= 0
    assert np.array_equal(b, want)
