"""Host-side product checks (no GPU): decomposer parity with the oracle and
the reference's hand-traced plans, plan lowering executed by the oracle,
estimator golden vectors, error handling of the C-ABI."""
import glob
import json
import os

import numpy as np
import pytest

import eithot_amd as E
from oracle import pyoracle as O
from plan_sim import simulate

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CORPUS = "/root/reference/Benchmark/test_cases"


def test_golden_plans():
    for case in json.load(open(os.path.join(GOLD, "plans.json")))["cases"]:
        got = E.decompose(case["dims"], case["perm"], case["reorder"])
        assert got == [(tuple(p), tuple(d)) for p, d in case["stages"]], case


def test_decomposer_matches_oracle_random():
    rng = np.random.default_rng(3)
    for _ in range(2000):
        r = int(rng.integers(1, 9))
        dims = [int(x) for x in rng.integers(1, 70, size=r)]
        perm1 = [int(p) + 1 for p in rng.permutation(r)]
        for reorder in (0, 1, 2):
            assert E.decompose(dims, perm1, reorder) == O.decompose(dims, perm1, reorder)


@pytest.mark.skipif(not os.path.isdir(CORPUS), reason="reference corpus not present")
def test_decomposer_matches_oracle_on_reference_corpus():
    n = 0
    for f in sorted(glob.glob(os.path.join(CORPUS, "**", "*.txt"), recursive=True)):
        for line in open(f):
            x = line.split()
            if not x:
                continue
            r = (len(x) - 3) // 2
            dims, perm1 = [int(v) for v in x[:r]], [int(v) for v in x[r:2 * r]]
            for reorder in (0, 1, 2):
                assert E.decompose(dims, perm1, reorder) == O.decompose(dims, perm1, reorder), line
                n += 1
    assert n > 10000


def _random_case(rng, maxvol=300000):
    r = int(rng.integers(1, 7))
    dims = [int(x) for x in np.exp(rng.uniform(0, np.log(300), size=r)).astype(int) + 1]
    while np.prod(dims) > maxvol:
        dims[int(np.argmax(dims))] //= 2
    perm1 = [int(p) + 1 for p in rng.permutation(r)]
    return dims, perm1


@pytest.mark.parametrize("lds", [0, 2048, 16384])
def test_plans_execute_correctly_on_oracle(lds):
    # small LDS budgets force column/row Catanzaro passes and staged passes
    rng = np.random.default_rng(lds + 1)
    seen = set()
    for _ in range(150):
        dims, perm1 = _random_case(rng)
        elem = int(rng.choice([1, 2, 4, 8]))
        vol = int(np.prod(dims))
        a = O.fill(vol, elem, 5, 1)
        want = O.permute(a, dims, [p - 1 for p in perm1], elem)
        ops, scratch = E.plan_describe(dims, perm1, elem, 1, 1.0, lds)
        seen |= {(o["kind"], o["map"], o["staged"]) for o in ops}
        b = a.copy()
        simulate(b, ops)
        assert np.array_equal(b, want), (dims, perm1, elem)
        assert scratch <= max(vol * elem, 1)
    if lds == 2048:
        assert (1, 0, 1) in seen  # some staged tile pass was exercised


def test_staged_catanzaro_plans_execute_correctly():
    # a scratch budget below one matrix rules out staged tiles, so tall
    # Catanzaro columns and long rows go through scratch
    rng = np.random.default_rng(31)
    seen = set()
    for _ in range(120):
        R, C = (int(x) for x in rng.integers(2, 600, size=2))
        B = int(rng.integers(1, 3))
        elem = int(rng.choice([1, 2, 4, 8]))
        dims, perm1 = [C, R, B], [2, 1, 3]
        ops, scratch = E.plan_describe(dims, perm1, elem, 1, 0.3, 1024)
        seen |= {(o["kind"], o["map"], o["staged"]) for o in ops}
        vol = B * R * C
        a = O.fill(vol, elem, 12, 1)
        b = a.copy()
        simulate(b, ops)
        assert np.array_equal(b, O.permute(a, dims, [1, 0, 2], elem)), (dims, elem, ops)
        assert scratch <= 0.3 * vol * elem + 1
    assert {(3, 5, 1), (3, 6, 1)} & seen and {(2, 3, 1), (2, 4, 1)} & seen, seen


NARROW_CASES = [([9, 271, 362], [3, 2, 1], 8, 0.05, 0), ([3, 100003], [2, 1], 8, 0.05, 0),
                ([100003, 3], [2, 1], 8, 0.05, 0), ([5, 60013], [2, 1], 4, 0.1, 0),
                ([2, 50021, 2], [2, 1, 3], 8, 0.05, 0), ([7, 4001], [2, 1], 8, 0.05, 4096),
                ([3, 200003], [2, 1], 8, 0.05, 0)]


@pytest.mark.parametrize("dims,perm1,elem,alpha,lds", NARROW_CASES)
def test_narrow_slice_plans_execute_correctly(dims, perm1, elem, alpha, lds):
    # rows too long for LDS at an aligned slice and for the scratch budget as
    # whole units: the row pass moves narrow slices of each unit through LDS
    # (cycle following, which would move the large cells instead, is off)
    old = E.set_tuning("cycle_min_cell", 0)
    try:
        ops, scratch = E.plan_describe(dims, perm1, elem, 1, alpha, lds)
    finally:
        E.set_tuning("cycle_min_cell", old)
    narrow = [o for o in ops if o["kind"] == 3 and o["us"] < o["U"]]
    assert narrow, ops
    # staged only when 1-byte slices of a row exceed LDS
    assert all(bool(o["staged"]) == (o["C"] > (lds or 160 * 1024)) for o in narrow), ops
    vol = int(np.prod(dims))
    assert scratch <= alpha * vol * elem
    a = O.fill(vol, elem, 8, 1)
    b = a.copy()
    simulate(b, ops)
    assert np.array_equal(b, O.permute(a, dims, [p - 1 for p in perm1], elem))


def test_fused_block_plans_execute_correctly():
    # consecutive contiguous tile passes whose blocks nest become one block
    # permutation; its gather table must compose to the oracle's permute
    rng = np.random.default_rng(5)
    fused = 0
    for _ in range(300):
        r = int(rng.integers(3, 7))
        dims = [int(x) for x in rng.integers(2, 20, size=r)]
        perm1 = [int(p) + 1 for p in rng.permutation(r)]
        elem = int(rng.choice([1, 2, 4, 8]))
        ops, scratch = E.plan_describe(dims, perm1, elem, 1, 0.5)
        if not any(o["kind"] == 5 for o in ops):
            continue
        fused += 1
        vol = int(np.prod(dims))
        assert scratch <= 0.5 * vol * elem
        a = O.fill(vol, elem, 3, 1)
        b = a.copy()
        simulate(b, ops)
        assert np.array_equal(b, O.permute(a, dims, [p - 1 for p in perm1], elem)), (dims, perm1)
        for o in ops:
            if o["kind"] == 5:
                assert sorted(o["table"].tolist()) == list(range(o["R"]))
    assert fused > 30


def test_regroup_plans_execute_correctly():
    # cell pass / tile pass / cell pass plans, checked against the oracle's
    # permute; they must differ from the decomposer's plan in many cases
    rng = np.random.default_rng(17)
    differ = 0
    for _ in range(200):
        r = int(rng.integers(3, 7))
        dims = [int(x) for x in rng.integers(2, 24, size=r)]
        perm1 = [int(p) + 1 for p in rng.permutation(r)]
        elem = int(rng.choice([1, 2, 4, 8]))
        alpha = float(rng.choice([0.05, 0.5]))
        ops, scratch = E.plan_describe(dims, perm1, elem, 1, alpha)
        old = E.set_tuning("regroup", 0)
        try:
            ops0, _ = E.plan_describe(dims, perm1, elem, 1, alpha)
        finally:
            E.set_tuning("regroup", old)
        key = lambda o: (o["kind"], o["B"], o["R"], o["C"], o["U"], o["us"])
        if [key(o) for o in ops] == [key(o) for o in ops0]:
            continue
        differ += 1
        vol = int(np.prod(dims))
        assert scratch <= max(alpha * vol * elem, 1)
        a = O.fill(vol, elem, 21, 1)
        b = a.copy()
        simulate(b, ops)
        assert np.array_equal(b, O.permute(a, dims, [p - 1 for p in perm1], elem)), (dims, perm1)
    assert differ > 10


@pytest.mark.parametrize("dims,perm1,elem", [
    ([1, 1, 1], [1, 2, 3], 4), ([1, 5, 1, 7], [4, 3, 2, 1], 4), ([3, 1, 4], [3, 2, 1], 2),
    ([2, 3, 4, 5], [1, 2, 4, 3], 8), ([2] * 12, list(range(12, 0, -1)), 4),
    ([3, 2] * 6, list(range(12, 0, -1)), 1), ([64, 2, 3], [1, 3, 2], 4),
    ([8, 1, 1, 16], [4, 2, 3, 1], 2), ([5, 7], [1, 2], 4)])
def test_plan_edge_shapes(dims, perm1, elem):
    # unit dims, identity, a leading dim that stays, rank 12, with and without
    # the sub-tensor partition, at tight and loose budgets
    vol = int(np.prod(dims))
    for alpha in (0.01, 0.05, 1.0):
        for nsub in (1, 4):
            ops, _ = E.plan_describe(dims, perm1, elem, nsub, alpha)
            a = O.fill(vol, elem, 2, 1)
            b = a.copy()
            simulate(b, ops)
            assert np.array_equal(b, O.permute(a, dims, [p - 1 for p in perm1], elem)), (alpha, nsub)


COMPOSITE_CASES = [([3, 14, 2, 15, 58, 20], [5, 2, 6, 1, 3, 4], 4),
                   ([12, 2, 25, 14, 53, 3], [3, 6, 4, 1, 5, 2], 8),
                   ([12, 9, 14, 28, 53], [2, 1, 5, 4, 3], 8),
                   ([51, 6, 10, 10, 4], [1, 3, 2, 5, 4], 8),
                   ([2, 39, 6, 2, 6, 13], [4, 5, 6, 3, 1, 2], 4),
                   ([46, 20, 52, 26], [4, 2, 3, 1], 8)]


@pytest.mark.parametrize("dims,perm1,elem", COMPOSITE_CASES)
def test_composite_cycle_plans_execute_correctly(dims, perm1, elem):
    # consecutive whole-cell passes composed into one cycle-following pass
    ops, scratch = E.plan_describe(dims, perm1, elem, 1, 0.2)
    assert any(o["kind"] == 4 and len(o["maps"]) > 1 for o in ops), ops
    vol = int(np.prod(dims))
    assert scratch <= 0.2 * vol * elem
    a = O.fill(vol, elem, 13, 1)
    b = a.copy()
    simulate(b, ops)
    assert np.array_equal(b, O.permute(a, dims, [p - 1 for p in perm1], elem))


def test_partition_plan_executes_correctly():
    for dims, perm1, nsub in (([152, 150], [2, 1], 4), ([6, 400, 3], [3, 2, 1], 8),
                              ([96, 40, 5], [2, 3, 1], 4)):
        vol = int(np.prod(dims))
        ops, _ = E.plan_describe(dims, perm1, 4, nsub, 1.0)
        assert any(o["stage"] == -1 for o in ops), "partition/join passes expected"
        a = O.fill(vol, 4, 6, 1)
        b = a.copy()
        simulate(b, ops)
        assert np.array_equal(b, O.permute(a, dims, [p - 1 for p in perm1], 4))


RAGGED_CASES = [([3, 100003], [2, 1], 4, 4), ([100003, 3], [2, 1], 4, 4),
                ([7, 1009, 5], [3, 1, 2], 2, 3), ([7, 1009, 5], [2, 3, 1], 1, 7),
                ([5, 6, 331, 4], [4, 3, 1, 2], 8, 5), ([331, 2, 3], [3, 1, 2], 4, 64),
                ([2, 4, 67, 3], [1, 3, 2, 4], 4, 10), ([90, 45], [2, 1], 8, 4),
                ([41], [1], 4, 3), ([2, 163, 5, 2], [1, 3, 2, 4], 4, 7)]


@pytest.mark.parametrize("dims,perm1,elem,nsub", RAGGED_CASES)
def test_ragged_partition_plans(dims, perm1, elem, nsub):
    # the requested sub-tensor count holds for split dims it does not divide:
    # nsub - 1 equal sub-tensors and a ragged last one; the plan run on the
    # host equals the oracle's permutation, and needs no padding headroom
    vol = int(np.prod(dims))
    n, d, q, r = E.plan_partition(dims, perm1, elem, nsub, 1.0)
    big = int(np.argmax(dims))
    if dims[big] <= 32 or perm1 == list(range(1, len(dims) + 1)):
        assert n == 1
        return
    assert n == min(nsub, dims[big]) and d == big
    assert (n - 1) * q + r == dims[big] and q >= 1 and r >= 1
    ops, _ = E.plan_describe(dims, perm1, elem, nsub, 1.0)
    assert any(o["stage"] == -1 for o in ops)
    assert max(o["base"] + o["B"] * o["R"] * o["C"] * o["U"] for o in ops) <= vol * elem
    a = O.fill(vol, elem, 17, 1)
    b = a.copy()
    simulate(b, ops)
    assert np.array_equal(b, O.permute(a, dims, [p - 1 for p in perm1], elem))


def test_ragged_partition_random():
    rng = np.random.default_rng(31)
    for _ in range(150):
        r = int(rng.integers(1, 6))
        dims = [int(x) for x in rng.integers(1, 9, size=r)]
        dims[int(rng.integers(r))] = int(rng.integers(33, 400))
        perm = [int(p) + 1 for p in rng.permutation(r)]
        elem = int(rng.choice([1, 2, 4, 8]))
        nsub = int(rng.integers(2, 70))
        vol = int(np.prod(dims))
        ops, _ = E.plan_describe(dims, perm, elem, nsub, 1.0)
        a = O.fill(vol, elem, 18, 1)
        b = a.copy()
        simulate(b, ops)
        assert np.array_equal(b, O.permute(a, dims, [p - 1 for p in perm], elem)), (dims, perm, nsub)


def test_benchmark_shape_plans():
    # (2,4096,1024,2) perm 1 3 2 4: at most the three Catanzaro passes of the reference's lowering,
    # scratch within 5 % of 64 MiB
    ops, scratch = E.plan_describe([2, 4096, 1024, 2], [1, 3, 2, 4], 4, 1, 0.05)
    assert len(ops) <= 3 and not any(o["staged"] for o in ops)
    assert scratch <= 0.05 * 2 * 4096 * 1024 * 2 * 4
    # without cycle following and fusion: only tile and Catanzaro passes, all
    # in LDS
    old = {k: E.set_tuning(k, 0) for k in ("cycle_min_cell", "fuse_max_block")}
    try:
        ops, scratch = E.plan_describe([2, 4096, 1024, 2], [1, 3, 2, 4], 4, 1, 0.05)
    finally:
        for k, v in old.items():
            E.set_tuning(k, v)
    assert all(o["kind"] in (1, 2, 3) and not o["staged"] for o in ops), ops
    assert scratch == 0
    # (32,48,64,40,36,8) perm 6 4 2 5 1 3: two passes (a composed cycle pass and a segmented tile)
    # where the reference runs five stages, none staged, scratch within 5 %
    # of 4.53 GB
    ops, scratch = E.plan_describe([32, 48, 64, 40, 36, 8], [6, 4, 2, 5, 1, 3], 4, 1, 0.05)
    assert len(ops) <= 2 and all(o["kind"] in (1, 4, 5, 6) and not o["staged"] for o in ops)
    assert scratch <= 0.05 * 32 * 48 * 64 * 40 * 36 * 8 * 4
    # (2,163840,65536,2) perm 1 3 2 4: scratch within 5 % of 171.8 GB
    ops, scratch = E.plan_describe([2, 163840, 65536, 2], [1, 3, 2, 4], 4, 1, 0.05)
    assert 0 < scratch <= 0.05 * 2 * 163840 * 65536 * 2 * 4


def test_estimator_golden_exp2():
    doc = json.load(open(os.path.join(GOLD, "estimator_exp2.json")))
    for c in doc["cases"]:
        e = E.estimate_reference(c["dims"], c["perm"], 4, 0.01)
        assert round(e["init_bytes"] / 1e6, 5) == pytest.approx(c["init_mb"], abs=1e-5), c
        assert e["nb"] == c["nb"], c
        assert e["alpha"] == pytest.approx(c["alpha"]), c
        assert e["printed_new_ratio"] == (e["init_bytes"] > 0)
        if e["printed_new_ratio"]:
            assert e["new_ratio"] == pytest.approx(doc["printed_new_ratio"])
    e = E.estimate_reference([7248, 7248], [2, 1], 4, 0.01)
    assert e["decomposed"]
    assert round(e["final_bytes"] / 1e6, 5) == doc["case1_final"]["final_mb"]
    assert e["final_ratio"] == pytest.approx(doc["case1_final"]["ratio"])


def test_native_estimate_within_alpha():
    rng = np.random.default_rng(9)
    for _ in range(200):
        dims, perm1 = _random_case(rng, 10 ** 7)
        for alpha in (0.01, 0.05, 0.5):
            try:
                n, need, extra = E.estimate(dims, perm1, 4, alpha)
            except E.EithotError as err:
                assert err.code == -3  # the plan cannot fit alpha
                continue
            vol = int(np.prod(dims))
            assert extra <= max(alpha * vol * 4, 1) + 1
            assert need == pytest.approx(extra / (vol * 4))


@pytest.mark.parametrize("dims,perm,elem,alpha,code", [
    ([4, 4], [1, 1], 4, 0.05, -1),      # not a permutation
    ([4, 4], [0, 1], 4, 0.05, -1),      # 0 is not 1-based
    ([4, 4, 4], [1, 2, 4], 4, 0.05, -1),  # out of range
    ([0, 4], [2, 1], 4, 0.05, -1),      # empty dim
    ([4, -3], [2, 1], 4, 0.05, -1),     # negative dim
    ([4, 4], [2, 1], 3, 0.05, -2),      # unsupported element size
    ([4, 4], [2, 1], 16, 0.05, -2),
    ([4, 4], [2, 1], 4, 0.0, -1),       # alpha must be positive
])
def test_transpose_rejects_invalid_input_before_touching_the_device(dims, perm, elem, alpha, code):
    import ctypes
    fake = ctypes.c_void_p(1 << 20)  # never dereferenced: validation fails first
    rc = E.lib().eithot_transpose(fake, E.i64(dims), E.i32(perm), len(dims), elem, 1, alpha)
    assert rc == code
    assert E.lib().eithot_last_error()


def test_reference_signature_rejects_bad_rank():
    import ctypes
    rc = E.lib().eithot_inplace_transpose(ctypes.c_void_p(1 << 20), 0, 0, E.i32([1]), E.i32([0]),
                                          4, 1, 0.05)
    assert rc == -1


def _row_stride_bound(rowbytes, us):
    def stride(g):
        g = max(g, 4)
        rs = (rowbytes + g - 1) // g * g
        return rs + g if (rs // g) % 2 == 0 else rs
    w = 1
    while w < 16 and us % (w * 2) == 0:
        w *= 2
    return max(stride(1 << k) for k in range(w.bit_length()))


@pytest.mark.parametrize("lds", [0, 8192, 65536])
def test_lds_footprints_fit_the_workgroup_limit(lds):
    # every LDS pass, including the padded tile rows, fits the LDS budget
    limit = lds or 160 * 1024
    rng = np.random.default_rng(lds + 77)
    for _ in range(300):
        dims, perm1 = _random_case(rng, 4_000_000)
        elem = int(rng.choice([1, 2, 4, 8]))
        ops, _ = E.plan_describe(dims, perm1, elem, 1, 1.0, lds)
        for o in ops:
            if o["staged"] or o["kind"] == 4:  # cycle following uses no LDS
                continue
            if o["kind"] == 1:
                need = o["R"] * _row_stride_bound(o["C"] * o["us"], o["us"])
            elif o["kind"] == 2:
                need = o["R"] * o["W"] * o["us"]
            else:
                need = o["C"] * o["us"]
            assert need <= limit, (dims, perm1, elem, o)


def test_catanzaro_forced_plans():
    # the GPU suite's Catanzaro cases plan the passes it asserts, and those
    # plans are correct when run on the host
    from test_gpu_parity import CATANZARO_CASES
    old = {k: E.set_tuning(k, v) for k, v in (("lower_mask", 2), ("regroup", 0))}
    try:
        for dims, perm, elem, alpha, want_ops in CATANZARO_CASES:
            ops, _ = E.plan_describe(dims, perm, elem, 1, alpha)
            assert {(o["kind"], o["map"], o["staged"]) for o in ops} == want_ops, dims
            vol = int(np.prod(dims))
            a = O.fill(vol, elem, 3, 1)
            b = a.copy()
            simulate(b, ops)
            assert np.array_equal(b, O.permute(a, dims, [p - 1 for p in perm], elem)), dims
    finally:
        for k, v in old.items():
            E.set_tuning(k, v)


def test_lower_mask_knob():
    # a mask that leaves no way to lower a transpose fails planning; an
    # out-of-range mask is refused
    old = E.set_tuning("lower_mask", 1)  # LDS tile only
    try:
        with pytest.raises(E.EithotError):
            E.plan_describe([100003, 3], [2, 1], 4, 1, 0.05)
        assert E.plan_describe([64, 32], [2, 1], 4, 1, 0.05)[0][0]["kind"] == 1
    finally:
        E.set_tuning("lower_mask", old)
    for bad in (0, 32, -1):
        with pytest.raises(E.EithotError):
            E.set_tuning("lower_mask", bad)
