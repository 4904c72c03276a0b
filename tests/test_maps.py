"""The product's index maps and magic division (eithot_amd/csrc/maps.hpp),
evaluated on the host, against the reference's functors restated in the
oracle (ITNT/src/equations.h) and against exact integer division."""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "_build", "libmaps_host.so")


@pytest.fixture(scope="module")
def mh():
    src = os.path.join(HERE, "maps_harness.cpp")
    hdr = os.path.join(HERE, "..", "eithot_amd", "csrc", "maps.hpp")
    if not os.path.exists(SO) or os.path.getmtime(SO) < max(os.path.getmtime(src),
                                                             os.path.getmtime(hdr)):
        os.makedirs(os.path.dirname(SO), exist_ok=True)
        # parallel test workers may build at once: each writes its own file
        # and renames it into place atomically
        tmp = f"{SO}.{os.getpid()}.tmp"
        subprocess.check_call(["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-arch=gfx950", "-O1",
                               "-std=c++17", "-fPIC", "-shared", "-o", tmp, src])
        os.replace(tmp, SO)
    L = ctypes.CDLL(SO)
    L.mh_eval.restype = ctypes.c_int64
    L.mh_eval.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_int64] * 9
    L.mh_div64.restype = ctypes.c_uint64
    L.mh_div64.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    L.mh_div32.restype = ctypes.c_uint32
    L.mh_div32.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    L.mh_odometer.restype = ctypes.c_int64
    L.mh_odometer.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                              ctypes.c_int]
    return L


def consts(m, n):
    c = math.gcd(m, n)
    a, b = m // c, n // c
    k = pow(a, -1, b) if b > 1 else 0
    binv = pow(b, -1, a) if a > 1 else 0
    return c, a, b, k, binv


SHAPES = [(m, n) for m in (1, 2, 3, 4, 6, 8, 12, 17, 32) for n in (1, 2, 5, 8, 12, 18, 32, 45)
          if m <= n] + [(1024, 4096), (64, 11520), (32, 552960), (65536, 163840),
                        # prime row lengths of the served workload: b * b and a * b far
                        # beyond 32 bits, reduced mod b in 64 bits by 32-bit maps
                        (20, 16060853), (141, 602891), (56, 2578133), (48, 2147483)]


@pytest.mark.parametrize("m,n", SHAPES)
def test_maps_equal_reference_functors(mh, m, n):
    c, a, b, k, binv = consts(m, n)
    rng = np.random.default_rng(m * 7919 + n)
    if m * n <= 4096:
        pts = [(i, j) for i in range(m) for j in range(n)]
    else:
        pts = [(int(i), int(j)) for i, j in zip(rng.integers(0, m, 3000), rng.integers(0, n, 3000))]
    # 32-bit maps are used where the planner allows them: column maps while
    # m*m + n and a*a stay below 2^31, row maps while n + 2m does
    col32 = m * m + n < (1 << 31) and a * a < (1 << 31)
    row32 = n + 2 * m < (1 << 31)
    for wide in (0, 1):
        for i, j in pts:
            ev = lambda w: mh.mh_eval(w, wide, m, n, c, a, b, k, binv, i, j)  # noqa: E731
            if wide or col32:
                assert ev(1) == O.fn_c2r_rotate(m, b, i, j)
                assert ev(2) == O.fn_r2c_rotate(m, b, i, j)
                assert ev(3) == O.fn_c2r_col_shuffle(m, n, c, i, j)
                assert ev(4) == O.fn_r2c_col_shuffle(a, c, m, binv, i, j)
            if wide or row32:
                assert ev(5) == O.fn_c2r_row_shuffle(m, n, c, k, i, j)
                # rowd is the inverse of the reference's row shuffle gather
                assert O.fn_c2r_row_shuffle(m, n, c, k, i, ev(6)) == j


def test_magic_division(mh):
    rng = np.random.default_rng(1)
    ds = [1, 2, 3, 5, 7, 10, 48, 96, 1000, 4096, 65535, 65536, 65537, 163840, 327680,
          (1 << 31) - 1, (1 << 31), (1 << 32) - 1]
    for d in ds:
        xs = [0, 1, d - 1, d, d + 1, (1 << 32) - 1] + [int(x) for x in rng.integers(0, 1 << 32, 300)]
        for x in xs:
            x &= (1 << 32) - 1
            assert mh.mh_div32(d, x) == x // d, (d, x)
    for d in ds + [(1 << 40) + 3, (1 << 62) + 1]:
        xs = [0, 1, d - 1, d, (1 << 64) - 1] + [int(x) for x in rng.integers(0, 1 << 63, 300, dtype=np.int64)]
        for x in xs:
            x &= (1 << 64) - 1
            assert mh.mh_div64(d, x) == x // d, (d, x)


def test_digit_odometer(mh):
    # the fused block pass steps a cell's source digit by digit: every step
    # equals the direct digit sum, for radices and strides of digit
    # permutations (nested transposes) and arbitrary ones
    import numpy as np
    rng = np.random.default_rng(11)
    cases = [([16], [1]), ([3, 5], [5, 1]), ([2, 2, 2, 2, 2, 2, 2, 2], [128, 64, 32, 16, 8, 4, 2, 1]),
             ([7, 1, 9], [9, 63, 1]), ([126], [1])]
    for _ in range(40):
        nd = int(rng.integers(1, 9))
        rad = [int(x) for x in rng.integers(1, 6, size=nd)]
        cases.append((rad, [int(x) for x in rng.integers(0, 1000, size=nd)]))
    for rad, st in cases:
        r = (ctypes.c_uint32 * 8)(*rad)
        t = (ctypes.c_uint32 * 8)(*st)
        assert mh.mh_odometer(r, t, len(rad)) == -1, (rad, st)
