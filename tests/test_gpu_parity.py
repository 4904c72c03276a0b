"""GPU parity: the HIP path through the C-ABI equals the oracle, bit for bit."""
import itertools

import numpy as np
import pytest

from gpu_util import assert_bitexact, run_case

pytestmark = pytest.mark.gpu


def test_known_answer_2x4096x1024x2(torch_cuda):
    # (2,4096,1024,2) perm 1 3 2 4, v(i) = i: the reference's own check
    got, want, peak = run_case(torch_cuda, [2, 4096, 1024, 2], [1, 3, 2, 4], 4, mode=0)
    assert_bitexact(got, want, "(2,4096,1024,2) perm 1 3 2 4")
    assert peak <= 0.05 * got.nbytes


@pytest.mark.parametrize("elem", [1, 2, 4, 8])
def test_all_perms_rank4_small(torch_cuda, elem):
    dims = [3, 5, 4, 7]
    for perm in itertools.permutations([1, 2, 3, 4]):
        got, want, _ = run_case(torch_cuda, dims, list(perm), elem, seed=2)
        assert_bitexact(got, want, f"perm {perm}")


# Catanzaro passes (kind, map, staged) expected when only that lowering is
# allowed: C2R = [rotate] -> row gather -> column shuffle, R2C = column
# shuffle inverse -> row scatter (gather form) -> [rotate inverse]
ROT, ROTINV, COLSH, COLSHINV, ROWG, ROWD = 1, 2, 3, 4, 5, 6
CATANZARO_CASES = [
    # gcd > 1 (rotate passes), columns and rows in LDS
    ([384, 256], [2, 1], 4, 0.05, {(2, ROT, 0), (3, ROWG, 0), (2, COLSH, 0)}),
    ([256, 384], [2, 1], 4, 0.05, {(2, COLSHINV, 0), (3, ROWD, 0), (2, ROTINV, 0)}),
    ([2, 640, 480, 2], [1, 3, 2, 4], 4, 0.05, {(2, ROT, 0), (3, ROWG, 0), (2, COLSH, 0)}),
    ([1024, 4098], [2, 1], 4, 0.05, {(2, COLSHINV, 0), (3, ROWD, 0), (2, ROTINV, 0)}),
    # gcd == 1: no rotate
    ([1000, 999], [2, 1], 1, 0.05, {(3, ROWG, 0), (2, COLSH, 0)}),
    ([999, 1000], [2, 1], 8, 0.05, {(2, COLSHINV, 0), (3, ROWD, 0)}),
    # columns of 3000 one-byte units exceed LDS: staged column passes,
    # rotate included
    ([3002, 3000], [2, 1], 1, 1.0, {(2, ROT, 1), (3, ROWG, 0), (2, COLSH, 1)}),
    ([3000, 3002], [2, 1], 1, 1.0, {(2, COLSHINV, 1), (3, ROWD, 0), (2, ROTINV, 1)}),
    # rows of 100003 units exceed LDS: staged row gather / scatter
    ([100003, 3], [2, 1], 4, 1.0, {(3, ROWG, 1), (2, COLSH, 0)}),
    ([3, 100003], [2, 1], 4, 1.0, {(2, COLSHINV, 0), (3, ROWD, 1)}),
]


@pytest.mark.parametrize("dims,perm,elem,alpha,want_ops", CATANZARO_CASES)
def test_catanzaro_passes(torch_cuda, dims, perm, elem, alpha, want_ops):
    # only the Catanzaro lowering allowed (no tile, split, cycle or regroup),
    # so the column / row kernels run with every map, in LDS and staged
    import eithot_amd as E
    old = {k: E.set_tuning(k, v) for k, v in (("lower_mask", 2), ("regroup", 0))}
    try:
        ops, _ = E.plan_describe(dims, perm, elem, 1, alpha)
        got_ops = {(o["kind"], o["map"], o["staged"]) for o in ops}
        assert got_ops == want_ops, (dims, got_ops)
        got, want, peak = run_case(torch_cuda, dims, perm, elem, seed=3, alpha=alpha)
        assert_bitexact(got, want, f"catanzaro {dims} {perm} elem {elem}")
        assert peak <= max(alpha * got.nbytes, 1)
    finally:
        for k, v in old.items():
            E.set_tuning(k, v)


def test_staged_passes(torch_cuda):
    # rows longer than LDS force scratch-staged passes (alpha allows them)
    for dims, perm, elem in (([3, 100003], [2, 1], 4), ([100003, 3], [2, 1], 4),
                             ([7, 65537], [2, 1], 4), ([2, 3001, 1801, 2], [1, 3, 2, 4], 4),
                             # 18-byte units in matrices of 2 rows, then of 2 columns
                             ([18, 300001, 2], [1, 3, 2], 1), ([18, 2, 300001], [1, 3, 2], 1)):
        got, want, peak = run_case(torch_cuda, dims, perm, elem, seed=4, alpha=1.0)
        assert_bitexact(got, want, f"{dims}")
        assert peak <= 1.0 * got.nbytes


# Staged Catanzaro row passes whose rows exceed LDS, tiled through LDS in
# runs of units (m = grid rows, U = unit bytes): row gather (map 5) and its
# inverse (map 6), words of 1 to 16 bytes, groups of one row and of several
STAGED_ROW_CASES = [
    ([6, 100003, 20], [1, 3, 2], 4, 0.05),   # m 20, 24-byte units, one row per group
    ([6, 20, 100003], [1, 3, 2], 4, 0.05),
    ([15, 30011, 141], [1, 3, 2], 2, 0.05),  # m 141, 30-byte units (2-byte words)
    ([15, 141, 30011], [1, 3, 2], 2, 0.05),
    ([2, 40009, 7], [1, 3, 2], 8, 0.3),      # 16-byte units, two rows per group
    ([2, 7, 40009], [1, 3, 2], 8, 0.3),
    ([3, 50021, 11], [1, 3, 2], 2, 0.1),     # 6-byte units
    ([3, 11, 50021], [1, 3, 2], 2, 0.1),
    ([9, 70001, 5], [1, 3, 2], 1, 0.2),      # 9-byte units (1-byte words)
    ([12, 5, 70001], [1, 3, 2], 1, 0.2),
    ([3, 30011, 300], [1, 3, 2], 4, 0.05),   # m 300, runs of 8 units
]


def test_staged_row_tiles(torch_cuda):
    # at the tensor's own address and one element past it (rows and scratch
    # tiles then start at other offsets mod 16)
    import eithot_amd as E
    from oracle import pyoracle as O
    old = {k: E.set_tuning(k, v) for k, v in (("lower_mask", 2), ("regroup", 0))}
    try:
        for dims, perm, elem, alpha in STAGED_ROW_CASES:
            ops, _ = E.plan_describe(dims, perm, elem, 1, alpha)
            assert any(o["kind"] == 3 and o["staged"] and o["us"] == o["U"] for o in ops), (dims, ops)
            got, want, peak = run_case(torch_cuda, dims, perm, elem, seed=31, alpha=alpha)
            assert_bitexact(got, want, f"staged rows {dims} {perm} elem {elem}")
            assert peak <= alpha * got.nbytes
            vol = int(np.prod(dims))
            host = O.fill(vol + 1, elem, 32, 1)
            dev = torch_cuda.from_numpy(host.copy()).to("cuda:0")
            print(f"case offset 1 dims={dims} perm={perm} elem={elem}", flush=True)
            E.transpose(dev, 1, len(dims), dims, [p - 1 for p in perm], elem, 1, alpha)
            torch_cuda.cuda.synchronize()
            want = host.copy()
            want[elem:] = O.permute(np.ascontiguousarray(host[elem:]), dims, [p - 1 for p in perm], elem)
            assert_bitexact(dev.cpu().numpy(), want, f"staged rows offset {dims} {perm} elem {elem}")
    finally:
        for k, v in old.items():
            E.set_tuning(k, v)


def test_narrow_slice_row_passes(torch_cuda):
    # rows longer than LDS with a 5 % budget: each unit is moved in narrow
    # slices (1 or 2 bytes of an 8-byte unit) through LDS, or through scratch
    # when even 1-byte slices of a row exceed LDS ([3, 200003])
    import eithot_amd as E
    old = E.set_tuning("cycle_min_cell", 0)  # cycle following would take the large cells
    try:
        for dims, perm, elem in (([9, 271, 362], [3, 2, 1], 8), ([3, 100003], [2, 1], 8),
                                 ([100003, 3], [2, 1], 8), ([2, 50021, 2], [2, 1, 3], 8),
                                 ([3, 200003], [2, 1], 8)):
            got, want, peak = run_case(torch_cuda, dims, perm, elem, seed=6, alpha=0.05)
            assert_bitexact(got, want, f"{dims} {perm}")
            assert peak <= 0.05 * got.nbytes
    finally:
        E.set_tuning("cycle_min_cell", old)


def test_cycle_following_passes(torch_cuda):
    # tile transposes of large cells move whole cells along the cycles of the
    # transpose: whole cycles, cut cycles (boundary slots), partial chunks,
    # 16/8/4/2/1-byte words
    import eithot_amd as E
    cases = [([128, 8, 36, 5], [1, 3, 2, 4], 4),      # (5, 36x8 cells of 512 B), whole cycles
             ([1536, 64, 120], [1, 3, 2], 4),         # one 120x64 matrix of 6 KB cells
             ([96, 48, 320, 3], [1, 3, 2, 4], 8),     # 48x320 of 768 B
             ([999, 17, 29], [1, 3, 2], 1),           # odd cells: 1-byte words
             ([257, 40, 9], [1, 3, 2], 2),            # 514-byte cells, 2-byte words
             ([300, 40, 7, 2], [1, 3, 2, 4], 2),      # 600-byte cells, 8-byte words
             ([129, 33, 31], [1, 3, 2], 4),           # 516-byte cells, 4-byte words
             ([32, 150, 140, 3], [1, 3, 2, 4], 4)]    # 128-byte cells: 8 items per wave
    default = E.get_tuning("cycle_variant")
    E.set_tuning("cycle_word_cost", 0)  # cells of odd sizes take cycle passes too
    try:
        for variant in range(8):  # every kernel shape, whichever is the default
            E.set_tuning("cycle_variant", variant)
            for dims, perm, elem in cases:
                ops, _ = E.plan_describe(dims, perm, elem, 1, 0.05)
                assert any(o["kind"] == 4 for o in ops), (dims, ops)
                got, want, peak = run_case(torch_cuda, dims, perm, elem, seed=8, alpha=0.05)
                assert_bitexact(got, want, f"variant {variant} {dims} {perm} elem {elem}")
                assert peak <= 0.05 * got.nbytes
    finally:
        E.set_tuning("cycle_variant", default)
        E.set_tuning("cycle_word_cost", 1)


# cells that are not whole 16-byte words, sized so the shifted cycle kernel
# packs items of 3 to 64 lanes (ceil(cell / 16) + 2), 1 to 21 to a wave
LANE_GROUP_CASES = [([5, 33, 29], [1, 3, 2], 2),      # 10 B: 3 lanes, 21 items per wave
                       ([10, 33, 29], [1, 3, 2], 2),     # 20 B: 4 lanes
                       ([9, 41, 23], [1, 3, 2], 4),      # 36 B: 5 lanes
                       ([13, 47, 53], [1, 3, 2], 4),     # 52 B: 6 lanes
                       ([45, 29, 31], [1, 3, 2], 2),     # 90 B: 8 lanes
                       ([65, 40, 33], [1, 3, 2], 2),     # 130 B: 11 lanes, 5 items per wave
                       ([35, 41, 29], [1, 3, 2], 4),     # 140 B: 11 lanes
                       ([138, 23, 31], [1, 3, 2], 2),    # 276 B: 20 lanes
                       ([125, 37, 19], [1, 3, 2], 4),    # 500 B: 34 lanes, one item per wave
                       ([1100, 13, 11], [1, 3, 2], 1)]   # 1100 B: two groups of 64 lanes


def test_shifted_cycle_lane_groups(torch_cuda):
    # only cycle following allowed, down to 8-byte cells; at the tensor's
    # address and one element past it
    import eithot_amd as E
    from oracle import pyoracle as O
    knobs = (("lower_mask", 16), ("regroup", 0), ("cycle_min_cell", 8), ("cycle_word_cost", 0))
    old = {k: E.set_tuning(k, v) for k, v in knobs}
    try:
        for dims, perm, elem in LANE_GROUP_CASES:
            ops, _ = E.plan_describe(dims, perm, elem, 1, 0.05)
            assert [o["kind"] for o in ops] == [4] and ops[0]["U"] % 16, (dims, ops)
            got, want, _ = run_case(torch_cuda, dims, perm, elem, seed=33, alpha=0.05)
            assert_bitexact(got, want, f"shifted cycle {dims} {perm} elem {elem}")
            vol = int(np.prod(dims))
            host = O.fill(vol + 1, elem, 34, 1)
            dev = torch_cuda.from_numpy(host.copy()).to("cuda:0")
            print(f"case offset 1 dims={dims} perm={perm} elem={elem}", flush=True)
            E.transpose(dev, 1, len(dims), dims, [p - 1 for p in perm], elem, 1, 0.05)
            torch_cuda.cuda.synchronize()
            want = host.copy()
            want[elem:] = O.permute(np.ascontiguousarray(host[elem:]), dims, [p - 1 for p in perm], elem)
            assert_bitexact(dev.cpu().numpy(), want, f"shifted cycle offset {dims} {perm} elem {elem}")
    finally:
        for k, v in old.items():
            E.set_tuning(k, v)


def test_composite_cycle_passes(torch_cuda):
    # consecutive whole-cell passes composed into one cycle-following pass
    from test_planner import COMPOSITE_CASES
    for dims, perm, elem in COMPOSITE_CASES + [([8, 40, 48, 9, 32, 16], [5, 3, 6, 2, 4, 1], 4)]:
        got, want, peak = run_case(torch_cuda, dims, perm, elem, seed=10, alpha=0.2)
        assert_bitexact(got, want, f"{dims} {perm} elem {elem}")
        assert peak <= 0.2 * got.nbytes


def test_fused_block_passes(torch_cuda):
    # nested contiguous tile passes fused into one LDS block permutation
    # (padded and unpadded LDS layouts, 1/2/4/8-byte cells), and the same
    # tensors with fusion off
    import eithot_amd as E
    cases = [([12, 9, 21, 15, 17], [2, 5, 4, 1, 3], 2), ([48, 40, 8, 300], [3, 2, 1, 4], 4),
             ([6, 5, 7, 9, 11], [2, 4, 3, 1, 5], 8), ([10, 13, 18, 12, 23], [5, 3, 4, 2, 1], 1),
             ([2, 6, 20, 3], [1, 4, 3, 2], 8), ([13, 15, 10, 14, 10, 14], [4, 2, 3, 1, 5, 6], 1),
             ([32, 48, 64, 5, 6, 8], [6, 4, 2, 5, 1, 3], 4),
             # lone tiles of small cells as block passes: blocks of 126 and 8
             # bytes (not multiples of 16: dense LDS, 16-byte runs and stores
             # across blocks), a partial last workgroup, 2- and 3-byte cells
             ([63, 2, 1001, 3], [2, 1, 3, 4], 1), ([4, 2, 30011], [2, 1, 3], 1),
             ([5, 3, 7, 2001], [2, 1, 3, 4], 2), ([3, 9, 7, 1001], [1, 3, 2, 4], 1)]
    default = E.get_tuning("fuse_max_block")
    regroup = E.set_tuning("regroup", 0)  # keep the plans that carry block passes
    try:
        for fuse, digits in ((default, 1), (default, 0), (0, 1)):
            E.set_tuning("fuse_max_block", fuse)
            E.set_tuning("block_digits", digits)
            for dims, perm, elem in cases:
                ops, _ = E.plan_describe(dims, perm, elem, 1, 0.5)
                if fuse:
                    assert any(o["kind"] == 5 for o in ops), (dims, perm, ops)
                got, want, _ = run_case(torch_cuda, dims, perm, elem, seed=9, alpha=0.5)
                assert_bitexact(got, want, f"fuse {fuse} digits {digits} {dims} {perm} elem {elem}")
    finally:
        E.set_tuning("fuse_max_block", default)
        E.set_tuning("block_digits", 1)
        E.set_tuning("regroup", regroup)


REGROUP_CASES = [([15, 24, 20, 13, 25], [5, 3, 1, 2, 4], 4), ([30, 9, 2, 66, 63], [5, 4, 1, 3, 2], 4),
                 ([13, 13, 23, 12, 2, 22], [4, 2, 6, 5, 1, 3], 4),
                 ([10, 37, 6, 30, 27], [5, 4, 3, 2, 1], 1),
                 ([12, 15, 11, 11, 13, 10], [2, 5, 3, 6, 4, 1], 8),
                 ([35, 27, 5, 3, 3, 8], [6, 4, 2, 5, 1, 3], 4),
                 ([2, 8, 103, 21, 112], [5, 3, 2, 4, 1], 8),
                 ([6, 2, 30, 25, 17, 18], [3, 6, 5, 1, 4, 2], 2),
                 ([32, 48, 64, 5, 6, 8], [6, 4, 2, 5, 1, 3], 4)]


def test_regroup_plans(torch_cuda):
    # cell pass / tile pass / cell pass plans: each case plans differently
    # with the regroup family off
    import eithot_amd as E
    key = lambda o: (o["kind"], o["B"], o["R"], o["C"], o["U"], o["us"])
    for dims, perm, elem in REGROUP_CASES:
        ops, _ = E.plan_describe(dims, perm, elem, 1, 0.05)
        old = E.set_tuning("regroup", 0)
        try:
            ops0, _ = E.plan_describe(dims, perm, elem, 1, 0.05)
        finally:
            E.set_tuning("regroup", old)
        assert [key(o) for o in ops] != [key(o) for o in ops0], (dims, perm)
        got, want, peak = run_case(torch_cuda, dims, perm, elem, seed=12, alpha=0.05)
        assert_bitexact(got, want, f"regroup {dims} {perm} elem {elem}")
        assert peak <= max(0.05 * got.nbytes, 1 << 20)


def test_random_high_order(torch_cuda):
    rng = np.random.default_rng(123)
    for _ in range(60):
        r = int(rng.integers(2, 7))
        dims = [int(x) for x in np.exp(rng.uniform(0, np.log(400), size=r)).astype(int) + 1]
        while np.prod(dims) > 2_000_000:
            dims[int(np.argmax(dims))] //= 2
        perm = [int(p) for p in rng.permutation(r) + 1]
        elem = int(rng.choice([1, 2, 4, 8]))
        got, want, peak = run_case(torch_cuda, dims, perm, elem, seed=int(rng.integers(1 << 30)),
                                   alpha=0.5)
        assert_bitexact(got, want, f"{dims} {perm} elem {elem}")
        assert peak <= 0.5 * got.nbytes


SEGTILE_CASES = [([14, 12, 14, 9, 9, 8], [5, 3, 2, 4, 1, 6], 8), ([2, 32, 4, 5, 32], [5, 2, 4, 1, 3], 4),
                 ([8, 12, 16, 16, 16], [5, 4, 1, 3, 2], 1), ([48, 32, 64], [3, 2, 1], 8),
                 ([10, 48, 64, 5, 8], [4, 3, 5, 1, 2], 1),
                 ([24, 40, 24, 24], [3, 2, 1, 4], 8), ([2, 24, 4, 8, 16, 16], [4, 6, 5, 3, 2, 1], 2),
                 ([16, 24, 5, 16, 16], [5, 4, 2, 1, 3], 8), ([3, 10, 8, 12, 8, 16], [6, 4, 3, 1, 2, 5], 8),
                 # X and Y split to fit LDS (fast factors of their last dims)
                 ([160, 64, 160], [3, 2, 1], 8), ([96, 128, 128], [3, 2, 1], 4),
                 ([48, 96, 2, 96], [4, 2, 3, 1], 2),
                 # the benchmark's two directions at a ninth of its size
                 ([32, 48, 16, 40, 9, 8], [6, 4, 2, 5, 1, 3], 4),
                 ([8, 40, 48, 9, 32, 16], [5, 3, 6, 2, 4, 1], 4)]


def test_segtile_plans(torch_cuda):
    # a cycle pass composed of both cell permutations plus a segmented tile
    # (tile first and tile last; 1- to 8-byte units; 16-byte words), and the
    # same tensors at a data pointer offset by one unit (narrower words)
    import eithot_amd as E
    for dims, perm, elem in SEGTILE_CASES:
        ops, _ = E.plan_describe(dims, perm, elem, 1, 0.05)
        assert any(o["kind"] == 6 for o in ops), (dims, perm, ops)
        got, want, peak = run_case(torch_cuda, dims, perm, elem, seed=14, alpha=0.05)
        assert_bitexact(got, want, f"segtile {dims} {perm} elem {elem}")
        assert peak <= max(0.05 * got.nbytes, 1 << 20)
    from oracle import pyoracle as O
    for dims, perm, elem in SEGTILE_CASES[::3]:
        vol = int(np.prod(dims))
        host = O.fill(vol + 1, elem, 15, 1)
        dev = torch_cuda.from_numpy(host.copy()).to("cuda:0")
        print(f"case offset 1 dims={dims} perm={perm} elem={elem}", flush=True)
        E.transpose(dev, 1, len(dims), dims, [p - 1 for p in perm], elem, 1, 0.05)
        torch_cuda.cuda.synchronize()
        got = dev.cpu().numpy()
        want = host.copy()
        want[elem:] = O.permute(np.ascontiguousarray(host[elem:]), dims, [p - 1 for p in perm], elem)
        assert_bitexact(got, want, f"segtile offset {dims} {perm} elem {elem}")


NARROW_CATANZARO = [([3, 300, 257], [1, 3, 2], 1), ([3, 257, 300], [1, 3, 2], 1),
                    ([9, 41, 1074], [1, 3, 2], 2), ([13, 445, 332], [1, 3, 2], 2),
                    ([18, 200, 131], [1, 3, 2], 1), ([5, 512, 100], [1, 3, 2], 2),
                    ([7, 1000, 999], [1, 3, 2], 1), ([6, 64, 96], [1, 3, 2], 4)]


def test_narrow_unit_catanzaro_passes(torch_cuda):
    # column and row passes over units that are not 16-byte words (3 to 26
    # bytes; panels that do not divide the grid): rows move as aligned
    # 16-byte chunks through LDS; at the tensor's own address and one
    # element past it (every row then starts at another offset mod 16)
    import eithot_amd as E
    from oracle import pyoracle as O
    old = {k: E.set_tuning(k, v) for k, v in (("lower_mask", 2), ("regroup", 0))}
    try:
        for dims, perm, elem in NARROW_CATANZARO:
            ops, _ = E.plan_describe(dims, perm, elem, 1, 0.05)
            assert {o["kind"] for o in ops} == {2, 3}, (dims, ops)
            got, want, _ = run_case(torch_cuda, dims, perm, elem, seed=16, alpha=0.05)
            assert_bitexact(got, want, f"narrow {dims} {perm} elem {elem}")
            vol = int(np.prod(dims))
            host = O.fill(vol + 1, elem, 17, 1)
            dev = torch_cuda.from_numpy(host.copy()).to("cuda:0")
            print(f"case offset 1 dims={dims} perm={perm} elem={elem}", flush=True)
            E.transpose(dev, 1, len(dims), dims, [p - 1 for p in perm], elem, 1, 0.05)
            torch_cuda.cuda.synchronize()
            want = host.copy()
            want[elem:] = O.permute(np.ascontiguousarray(host[elem:]), dims, [p - 1 for p in perm],
                                    elem)
            assert_bitexact(dev.cpu().numpy(), want, f"narrow offset {dims} {perm} elem {elem}")
    finally:
        for k, v in old.items():
            E.set_tuning(k, v)


SHIFTED_CYCLE_CASES = [([129, 60, 70], [1, 3, 2], 1),      # 129-byte cells, 1-byte words
                       ([257, 40, 30], [1, 3, 2], 2),      # 514-byte cells
                       ([300, 40, 17, 2], [1, 3, 2, 4], 2),  # 600-byte cells, 8-byte words
                       ([1273, 6, 35], [1, 3, 2], 4),      # 5092-byte cells: 6 item groups each
                       ([2547, 35, 6], [1, 3, 2], 2),
                       ([3, 14, 2, 15, 58, 20], [5, 2, 6, 1, 3, 4], 4)]  # composed maps


def test_shifted_cycle_passes(torch_cuda):
    # cycle following of cells that are not whole 16-byte words, moved in
    # aligned 16-byte chunks with a funnel shift: whole and cut cycles
    # (segments of 3 positions), at the tensor's address and one element past
    # it; every kernel variant other than the default moves such cells word
    # by word and is run once for comparison
    import eithot_amd as E
    from oracle import pyoracle as O
    regroup = E.set_tuning("regroup", 0)
    try:
        for lmax in (0, 3):
            E.set_tuning("cycle_lmax", lmax)
            for variant in (0, 4):
                E.set_tuning("cycle_variant", variant)
                for dims, perm, elem in SHIFTED_CYCLE_CASES:
                    ops, _ = E.plan_describe(dims, perm, elem, 1, 0.5)
                    assert any(o["kind"] == 4 and o["U"] % 16 for o in ops), (dims, ops)
                    got, want, _ = run_case(torch_cuda, dims, perm, elem, seed=18, alpha=0.5)
                    assert_bitexact(got, want, f"lmax {lmax} variant {variant} {dims} {perm}")
                    vol = int(np.prod(dims))
                    host = O.fill(vol + 1, elem, 19, 1)
                    dev = torch_cuda.from_numpy(host.copy()).to("cuda:0")
                    print(f"case offset 1 dims={dims} perm={perm} elem={elem}", flush=True)
                    E.transpose(dev, 1, len(dims), dims, [p - 1 for p in perm], elem, 1, 0.5)
                    torch_cuda.cuda.synchronize()
                    want = host.copy()
                    want[elem:] = O.permute(np.ascontiguousarray(host[elem:]), dims,
                                            [p - 1 for p in perm], elem)
                    assert_bitexact(dev.cpu().numpy(), want, f"offset lmax {lmax} {dims} {perm}")
    finally:
        E.set_tuning("cycle_lmax", 0)
        E.set_tuning("cycle_variant", 0)
        E.set_tuning("regroup", regroup)


DENSE_BLOCK_CASES = [([71, 30, 277, 4], [3, 4, 2, 1], 1),     # blocks of 19,667 and 33,240 bytes
                     ([67, 8, 54, 45, 2, 8], [3, 2, 6, 4, 5, 1], 1),  # 28,944-byte blocks
                     ([13, 4, 3, 63, 505], [3, 4, 2, 5, 1], 2),     # 9,828 two-byte cells
                     ([57, 137, 340], [2, 1, 3], 4),                # 7,809 four-byte cells
                     ([63, 2, 8661, 3], [2, 1, 3, 4], 1),           # 126-byte blocks (pattern)
                     ([4, 2, 38591, 9], [2, 1, 4, 3], 1),           # 2,664-byte blocks
                     ([58, 2, 12336], [3, 2, 1], 1),                # 29,812-byte blocks
                     ([6, 17, 2, 21875], [2, 3, 4, 1], 2)]          # 840-byte blocks (pattern)


def test_dense_block_passes(torch_cuda):
    # fused blocks whose size is not a multiple of 16 bytes, or at a data
    # pointer one element past an aligned one: the workgroup's run of blocks
    # sits in LDS at its address modulo 16, interior 16-byte chunks are
    # gathered whole (digit odometer, gather table, or per-thread pattern
    # offsets for small blocks) and the run's edge chunks word by word
    import eithot_amd as E
    from oracle import pyoracle as O
    try:
        for digits in (1, 0):
            E.set_tuning("block_digits", digits)
            for dims, perm, elem in DENSE_BLOCK_CASES:
                ops, _ = E.plan_describe(dims, perm, elem, 1, 0.05)
                assert any(o["kind"] == 5 for o in ops), (dims, perm, ops)
                got, want, _ = run_case(torch_cuda, dims, perm, elem, seed=20, alpha=0.05)
                assert_bitexact(got, want, f"digits {digits} {dims} {perm} elem {elem}")
                vol = int(np.prod(dims))
                host = O.fill(vol + 1, elem, 21, 1)
                dev = torch_cuda.from_numpy(host.copy()).to("cuda:0")
                print(f"case offset 1 digits={digits} dims={dims} perm={perm} elem={elem}", flush=True)
                E.transpose(dev, 1, len(dims), dims, [p - 1 for p in perm], elem, 1, 0.05)
                torch_cuda.cuda.synchronize()
                want = host.copy()
                want[elem:] = O.permute(np.ascontiguousarray(host[elem:]), dims, [p - 1 for p in perm],
                                        elem)
                assert_bitexact(dev.cpu().numpy(), want, f"offset digits {digits} {dims} {perm}")
    finally:
        E.set_tuning("block_digits", 1)


NARROW_TILE_CASES = [([10, 13403, 2], [2, 1, 3], 1),     # 134,030-byte matrices: one LDS copy
                     ([11, 9222, 2], [2, 1, 3], 1),      # 101,442 bytes: one copy
                     ([780, 139, 3], [2, 1, 3], 1),      # 108,420 bytes: one copy
                     ([6, 770, 24, 3], [1, 3, 2, 4], 1),  # 6-byte units, 110,880 bytes: one copy
                     ([2, 40732, 2], [2, 1, 3], 1),      # 81,464 bytes: two copies
                     ([97, 416, 5], [2, 1, 3], 2),       # 2-byte units, 80,704 bytes: two copies
                     ([4307, 16, 3], [2, 1, 3], 1),      # 68,912 bytes: two copies
                     ([9, 283, 2, 40], [1, 3, 2, 4], 2),  # 18-byte units in 2 x 283 matrices
                     ([64, 96, 5], [2, 1, 3], 1)]        # one-word cells that pack: not wide


def test_narrow_word_tiles(torch_cuda):
    # tile passes of 1- and 2-byte words: matrices run through LDS as aligned
    # 16-byte chunks, transposed between two LDS copies or gathered from one
    # copy through a digit odometer; at the tensor's own address and one
    # element past it (every matrix then starts at another offset mod 16)
    import eithot_amd as E
    from oracle import pyoracle as O
    for dims, perm, elem in NARROW_TILE_CASES:
        ops, _ = E.plan_describe(dims, perm, elem, 1, 0.05)
        tiles = [o for o in ops if o["kind"] == 1 and not o["staged"]]
        assert tiles and all(o["us"] == o["U"] and (o["U"] % 4) for o in tiles), (dims, ops)
        got, want, _ = run_case(torch_cuda, dims, perm, elem, seed=22, alpha=0.05)
        assert_bitexact(got, want, f"narrow tile {dims} {perm} elem {elem}")
        vol = int(np.prod(dims))
        host = O.fill(vol + 1, elem, 23, 1)
        dev = torch_cuda.from_numpy(host.copy()).to("cuda:0")
        print(f"case offset 1 dims={dims} perm={perm} elem={elem}", flush=True)
        E.transpose(dev, 1, len(dims), dims, [p - 1 for p in perm], elem, 1, 0.05)
        torch_cuda.cuda.synchronize()
        want = host.copy()
        want[elem:] = O.permute(np.ascontiguousarray(host[elem:]), dims, [p - 1 for p in perm], elem)
        assert_bitexact(dev.cpu().numpy(), want, f"narrow tile offset {dims} {perm} elem {elem}")
