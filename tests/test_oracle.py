"""Pins the oracle itself: each function against golden vectors, the
reference's own integer-math sources (oracle/_ref), or what the mathematics
fixes (numpy's transpose, closed forms, bijections, brute force)."""
import itertools
import json
import os

import numpy as np
import pytest

from oracle import pyoracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def np_reference(arr, dims, perm0):
    """numpy's own transpose: a dims-fastest-first tensor is a C-order array of
    reversed shape; output numpy axis a holds input dim perm0[r-1-a]."""
    r = len(dims)
    a = arr.reshape(list(reversed(dims)))
    axes = [r - 1 - perm0[r - 1 - ax] for ax in range(r)]
    return np.ascontiguousarray(np.transpose(a, axes)).reshape(-1)


@pytest.mark.parametrize("rank", [1, 2, 3, 4, 5])
def test_permute_matches_numpy_all_perms(rank):
    rng = np.random.default_rng(rank)
    dims = [int(x) for x in rng.integers(1, 6, size=rank)]
    vol = int(np.prod(dims))
    for elem, dt in ((1, np.uint8), (2, np.uint16), (4, np.uint32), (8, np.uint64)):
        arr = O.fill(vol, elem, 9, 1).view(dt)
        for perm in itertools.permutations(range(rank)):
            got = O.permute(arr, dims, list(perm), elem)
            assert np.array_equal(got, np_reference(arr, dims, list(perm))), (dims, perm, elem)


def test_permute_mt_matches_single_thread():
    rng = np.random.default_rng(7)
    for _ in range(30):
        r = int(rng.integers(1, 6))
        dims = [int(x) for x in rng.integers(1, 40, size=r)]
        perm = [int(p) for p in rng.permutation(r)]
        elem = int(rng.choice([1, 2, 4, 8]))
        arr = O.fill(int(np.prod(dims)), elem, 3, 1)
        a = O.permute(arr, dims, perm, elem)
        b = O.permute(arr, dims, perm, elem, nthreads=4)
        assert np.array_equal(a, b), (dims, perm, elem)
    # leading dims that stay in place (folded into the element), also past
    # the fold limit
    for dims, perm in (([2, 300, 500, 2], [0, 2, 1, 3]), ([3, 5, 70, 90], [0, 1, 3, 2]),
                       ([4000, 3, 7], [0, 2, 1]), ([2, 2, 2, 2, 9, 11], [0, 1, 2, 5, 4, 3]),
                       ([6, 9], [0, 1])):
        for elem in (1, 8):
            arr = O.fill(int(np.prod(dims)), elem, 4, 1)
            a = O.permute(arr, dims, perm, elem)
            assert np.array_equal(a, np_reference(arr.view({1: np.uint8, 8: np.uint64}[elem]),
                                                  dims, perm).view(np.uint8)), (dims, perm)
            b = O.permute(arr, dims, perm, elem, nthreads=3)
            assert np.array_equal(a, b), (dims, perm, elem)


def test_known_answer_closed_form():
    # v(i) = i and the closed-form destination (verify_inplace,
    # ITNT/src/tensor_util.cpp:112-142): out[dest(i)] == i
    dims, perm0 = [2, 64, 32, 2], [0, 2, 1, 3]
    vol = int(np.prod(dims))
    arr = np.arange(vol, dtype=np.uint32)
    out = O.permute(arr, dims, perm0, 4)
    for i in range(0, vol, 37):
        assert out[O.dest_index(i, dims, perm0)] == i


def test_generator_modes():
    a = O.fill(1000, 4, 5, 0).view(np.uint32)
    assert np.array_equal(a, np.arange(1000, dtype=np.uint32))
    b = O.fill(1000, 8, 5, 1).view(np.uint64)
    assert len(set(b.tolist())) == 1000
    c = O.fill(10, 8, 5, 1, first=990).view(np.uint64)
    assert np.array_equal(b[990:], c)


def test_extended_gcd_against_reference_build():
    rng = np.random.default_rng(11)
    R = O.ref_lib()
    import ctypes
    pairs = [(int(a), int(b)) for a, b in rng.integers(1, 1 << 20, size=(2000, 2))]
    pairs += [(1, 1), (7248, 7248), (1024, 8192), (65536, 163840), (1, 7), (7, 1)]
    for a, b in pairs:
        g, m = O.extended_gcd(a, b)
        assert g == np.gcd(a, b)
        if g == 1 and b > 1:
            assert (a * m) % b == 1  # Bezout: m is a^-1 mod b
        if R is not None:
            gg, mm = ctypes.c_int(), ctypes.c_int()
            R.ref_extended_gcd(a, b, ctypes.byref(gg), ctypes.byref(mm))
            assert (gg.value, mm.value) == (g, m), (a, b)


def test_reference_magic_division_is_exact():
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    import ctypes
    rng = np.random.default_rng(2)
    q = ctypes.c_uint()
    for d in [1, 2, 3, 5, 7, 48, 1000, 65535, 65536, 327680]:
        for x in list(rng.integers(0, 1 << 31, size=200)) + [0, 1, d, d - 1, (1 << 31) - 1]:
            assert R.ref_div32(d, int(x), ctypes.byref(q)) == 0
            assert q.value == int(x) // d


def test_decomposer_golden_plans():
    doc = json.load(open(os.path.join(GOLD, "plans.json")))
    for case in doc["cases"]:
        got = O.decompose(case["dims"], case["perm"], case["reorder"])
        want = [(tuple(p), tuple(d)) for p, d in case["stages"]]
        assert got == want, case


def _stage_shape_check(dims, perm1, stages):
    """Every stage is a 21/213/132/1324 block swap over the whole volume."""
    for lp, ld in stages:
        assert tuple(lp) in {(2, 1), (2, 1, 3), (1, 3, 2), (1, 3, 2, 4)}
        assert int(np.prod(ld)) == int(np.prod(dims))


def test_decomposer_invariants_random():
    rng = np.random.default_rng(5)
    for _ in range(500):
        r = int(rng.integers(1, 8))
        dims = [int(x) for x in rng.integers(1, 9, size=r)]
        perm1 = [int(p) + 1 for p in rng.permutation(r)]
        for reorder in (0, 1, 2):
            st = O.decompose(dims, perm1, reorder)
            assert len(st) <= r
            _stage_shape_check(dims, perm1, st)
            if perm1 == sorted(perm1):
                assert st == []


@pytest.mark.parametrize("m,n", [(m, n) for m in range(1, 25) for n in range(1, 25)])
def test_catanzaro_reference_functors_transpose(m, n):
    # The reference's C2R and R2C pass sequences (2dtranspose.cpp:128-183)
    # built from the restated functors equal the direct transpose.
    a = O.fill(m * n, 4, m * 131 + n, 1)
    want = a.copy()
    O.transpose2d_direct(want, 1, m, n, 4)
    for method in (0, 1):
        b = a.copy()
        O.transpose2d_catanzaro(b, 1, m, n, 4, method)
        assert np.array_equal(b, want), (m, n, method)


def test_functor_bijections():
    for m, n in [(4, 6), (6, 4), (7, 5), (12, 18), (1024, 8192)]:
        c = int(np.gcd(m, n))
        a, b = m // c, n // c
        g, k = O.extended_gcd(a, b)
        g, q = O.extended_gcd(b, a)
        cols = range(0, n, max(1, n // 7))
        for j in cols:
            for fn in (lambda i: O.fn_c2r_rotate(m, b, i, j),
                       lambda i: O.fn_c2r_col_shuffle(m, n, c, i, j),
                       lambda i: O.fn_r2c_rotate(m, b, i, j)):
                assert sorted(fn(i) for i in range(m)) == list(range(m))
        for i in range(0, m, max(1, m // 7)):
            assert sorted(O.fn_c2r_row_shuffle(m, n, c, k, i, j) for j in range(n)) == \
                list(range(n))


def test_simulate_reference_random_nd():
    rng = np.random.default_rng(17)
    for _ in range(120):
        r = int(rng.integers(2, 7))
        dims = [int(x) for x in np.exp(rng.uniform(0, np.log(60), size=r)).astype(int) + 1]
        while np.prod(dims) > 150000:
            dims[int(np.argmax(dims))] //= 2
        perm1 = [int(p) + 1 for p in rng.permutation(r)]
        elem = int(rng.choice([1, 2, 4, 8]))
        vol = int(np.prod(dims))
        a = O.fill(vol, elem, 21, 1)
        want = O.permute(a, dims, [p - 1 for p in perm1], elem)
        nb = int(rng.choice([1, 2, 4, 8]))
        alpha = float(rng.choice([0.05, 0.5, 4.0]))
        buf = np.zeros(2 * vol * elem, dtype=np.uint8)
        buf[:vol * elem] = a
        assert O.simulate_reference(buf, dims, perm1, elem, nb, alpha, capacity_elems=2 * vol) == 0
        assert np.array_equal(buf[:vol * elem], want), (dims, perm1, elem, nb, alpha)


def _exp2_cases():
    path = os.path.join(GOLD, "estimator_exp2.json")
    return [(c["dims"], c["perm"]) for c in json.load(open(path))["cases"]]


def test_simulate_reference_exp2_shapes_scaled():
    # the 57 reference-validated exp2 (dims, perm) pairs, dims scaled down
    for dims, perm1 in _exp2_cases():
        small = [max(2, d // 12) if len(dims) > 3 else max(2, d // 96) for d in dims]
        while np.prod(small) > 200000:
            small[int(np.argmax(small))] //= 2
        vol = int(np.prod(small))
        a = O.fill(vol, 4, 1, 0)
        want = O.permute(a, small, [p - 1 for p in perm1], 4)
        b = a.copy()
        assert O.simulate_reference(b, small, perm1, 4, 1, 0.01) == 0
        assert np.array_equal(b, want), (small, perm1)


def test_simulate_reference_partition_with_padding():
    # exp2 case 1 structure: the only exp2 case the reference partitions;
    # an indivisible dim exercises pre/post padding with the +headroom buffer
    # alpha 1.0 makes floor(32/alpha) = 32 so every case partitions
    for dims, nb in (([120, 96], 4), ([121, 97], 4), ([3, 1003, 5], 8)):
        perm1 = list(range(len(dims), 0, -1))
        vol = int(np.prod(dims))
        a = O.fill(vol, 4, 8, 1)
        want = O.permute(a, dims, [p - 1 for p in perm1], 4)
        cap = int(vol * 1.2) + 64
        buf = np.zeros(cap * 4, dtype=np.uint8)
        buf[:vol * 4] = a
        assert O.simulate_reference(buf, dims, perm1, 4, nb, 1.0, capacity_elems=cap) == 0
        assert np.array_equal(buf[:vol * 4], want), (dims, nb)
