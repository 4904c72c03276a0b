"""eithot_transpose_noalloc (scratch and pass tables in a static device array)
against eithot_transpose on the same inputs, bit for bit: two calls in a row,
the second on a smaller tensor, then plans that use the scratch (cut cycles,
a staged pass, a fused block table), and refusal of a plan too large for the
array."""
import ctypes

import numpy as np
import pytest

from gpu_util import assert_bitexact
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _noalloc():
    import eithot_amd as E
    return E.lib().eithot_transpose_noalloc


def _both(torch, dims, perm, elem, alpha, seed):
    """(no-alloc result, plain result, oracle) for one seeded tensor."""
    import eithot_amd as E
    from eithot_amd._lib import i32, i64
    f = _noalloc()
    vol = int(np.prod(dims))
    host = O.fill(vol, elem, seed, 1)
    a = torch.from_numpy(host.copy()).to("cuda:0")
    b = torch.from_numpy(host.copy()).to("cuda:0")
    print(f"no-alloc case dims={dims} perm={perm} elem={elem} alpha={alpha}", flush=True)
    rc = f(ctypes.c_void_p(a.data_ptr()), i64(dims), i32(perm), len(dims), elem, 1, alpha)
    assert rc == 0, E.lib().eithot_last_error()
    used = E.last_peak_extra_bytes()
    E.transpose_tensor(b, dims, perm, elem, 1, alpha)
    torch.cuda.synchronize()
    want = O.permute(host, dims, [p - 1 for p in perm], elem)
    return a.cpu().numpy(), b.cpu().numpy(), want, used


def test_noalloc_matches_plain_entry_point(torch_cuda):
    import eithot_amd as E
    cases = [
        # two calls in a row, the second smaller: a cycle pass with cut
        # cycles (boundary cells in the scratch) and a segmented tile, then a
        # cycle pass of whole cycles
        ([32, 48, 16, 40, 9, 8], [6, 4, 2, 5, 1, 3], 4, 0.05),
        ([128, 8, 36, 5], [1, 3, 2, 4], 4, 0.05),
        # the first shape again: its tables are already in the array
        ([32, 48, 16, 40, 9, 8], [6, 4, 2, 5, 1, 3], 4, 0.05),
        # a staged pass through the scratch, then a fused block table
        ([3, 100003], [2, 1], 4, 1.0),
        ([12, 9, 21, 15, 17], [2, 5, 4, 1, 3], 2, 0.5),
        ([2, 4096, 1024, 2], [1, 3, 2, 4], 4, 0.05),
    ]
    digits = E.set_tuning("block_digits", 0)  # fused blocks read a gather table
    regroup = E.set_tuning("regroup", 0)
    try:
        kinds = set()
        for k, (dims, perm, elem, alpha) in enumerate(cases):
            ops, extra = E.plan_describe(dims, perm, elem, 1, alpha)
            kinds |= {(o["kind"], o["staged"]) for o in ops}
            got, plain, want, used = _both(torch_cuda, dims, perm, elem, alpha, 40 + k)
            assert_bitexact(got, want, f"no-alloc {dims} {perm}")
            assert_bitexact(got, plain, f"no-alloc vs plain {dims} {perm}")
            assert 0 <= used <= 64 << 20
        assert {(4, 0), (5, 0), (1, 1)} <= kinds, kinds
    finally:
        E.set_tuning("block_digits", digits)
        E.set_tuning("regroup", regroup)


def test_noalloc_refuses_plans_beyond_its_array(torch_cuda):
    import eithot_amd as E
    from eithot_amd._lib import i32, i64
    f = _noalloc()
    # a staged tile whose footprint alone is 120 MB: refused before any launch
    dims = [3, 10000003]
    t = torch_cuda.zeros(3 * 10000003 * 4, dtype=torch_cuda.uint8, device="cuda:0")
    rc = f(ctypes.c_void_p(t.data_ptr()), i64(dims), i32([2, 1]), 2, 4, 1, 1.0)
    assert rc == -5, rc
    assert "static scratch" in E.lib().eithot_last_error().decode()
    assert int(t.count_nonzero()) == 0
