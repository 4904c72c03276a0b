"""GPU parity at the workloads' real sizes.

The device pattern fill and closed-form check (eithot_fill_pattern,
eithot_verify_pattern) are pinned against the oracle first; the full-size
tensors, too large to copy to the host and permute there, are then checked
with them: the forward result against the permuted pattern, a sample of
elements against the oracle's destination index, and the round trip.
"""
import numpy as np
import pytest

from gpu_util import assert_bitexact
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

NP = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


def _inverse(dims, perm):
    out = [dims[p - 1] for p in perm]
    inv = [0] * len(perm)
    for k, p in enumerate(perm):
        inv[p - 1] = k + 1
    return out, inv


def _free(torch):
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def test_device_checker_matches_oracle(torch_cuda):
    import eithot_amd as E
    torch = torch_cuda
    # the fill equals the oracle's generator byte for byte
    n = 100003
    for elem in (1, 2, 4, 8):
        for mode in (0, 1):
            for first in (0, 123457):
                t = torch.empty(n * elem, dtype=torch.uint8, device="cuda:0")
                print(f"fill elem={elem} mode={mode} first={first}", flush=True)
                E.fill_pattern(t, n, elem, 77, mode, first)
                torch.cuda.synchronize()
                assert_bitexact(t.cpu().numpy(), O.fill(n, elem, 77, mode, first),
                                f"fill elem {elem} mode {mode} first {first}")
    # the check counts 0 on the oracle's permutation and exactly the
    # elements planted wrong (on the host) otherwise
    rng = np.random.default_rng(5)
    for dims, perm in (([5, 7, 3, 11], [3, 1, 4, 2]), ([64, 3, 40], [2, 3, 1]),
                       ([2, 409, 97, 2], [1, 3, 2, 4])):
        vol = int(np.prod(dims))
        for elem in (1, 2, 4, 8):
            for mode in (0, 1):
                host = O.fill(vol, elem, 9, mode)
                out = O.permute(host, dims, [p - 1 for p in perm], elem)
                print(f"verify {dims} {perm} elem={elem} mode={mode}", flush=True)
                t = torch.from_numpy(out.copy()).to("cuda:0")
                assert E.verify_pattern(t, dims, perm, elem, 9, mode) == 0
                for planted in (1, 7, 64):
                    bad = out.copy()
                    idx = rng.choice(vol, size=planted, replace=False)
                    bad[idx * elem + int(rng.integers(elem))] ^= 0x5A
                    t = torch.from_numpy(bad).to("cuda:0")
                    assert E.verify_pattern(t, dims, perm, elem, 9, mode) == planted
                # the unpermuted input passes the identity permutation's check
                t = torch.from_numpy(host.copy()).to("cuda:0")
                assert E.verify_pattern(t, dims, list(range(1, len(dims) + 1)), elem, 9,
                                        mode) == 0


def _sampled_check(torch, buf, dims, perm, elem, seed, mode, count=4000):
    """Oracle destination index of sampled input elements vs the device bytes."""
    vol = int(np.prod(dims))
    rng = np.random.default_rng(seed + 1)
    src = rng.integers(0, vol, size=count, dtype=np.int64)
    dst = np.array([O.dest_index(int(i), dims, [p - 1 for p in perm]) for i in src],
                   dtype=np.int64)
    words = {1: torch.uint8, 2: torch.int16, 4: torch.int32, 8: torch.int64}[elem]
    got = buf.view(words)[torch.from_numpy(dst).to("cuda:0")].cpu().numpy().view(NP[elem])
    mask = (1 << (8 * elem)) - 1
    want = np.array([O.value(seed, int(i), mode) & mask for i in src], dtype=np.uint64)
    assert np.array_equal(got.astype(np.uint64), want), "sampled elements differ"


def _full_size(torch, dims, perm, elem, seed, alpha=0.05):
    import eithot_amd as E
    _free(torch)
    vol = int(np.prod(dims))
    nbytes = vol * elem
    buf = torch.empty(nbytes, dtype=torch.uint8, device="cuda:0")
    E.fill_pattern(buf, vol, elem, seed, 1)
    n_sub, _, _ = E.estimate(dims, perm, elem, alpha)  # sub-tensor count from the estimator
    print(f"full size {dims} perm {perm}: forward (n_sub {n_sub})", flush=True)
    E.transpose_tensor(buf, dims, perm, elem, n_sub, alpha)
    torch.cuda.synchronize()
    held = E.last_peak_extra_bytes()
    assert held <= alpha * nbytes, (held, nbytes)
    print("  check forward", flush=True)
    assert E.verify_pattern(buf, dims, perm, elem, seed, 1) == 0
    _sampled_check(torch, buf, dims, perm, elem, seed, 1)
    inv_dims, inv_perm = _inverse(dims, perm)
    print("  inverse", flush=True)
    E.transpose_tensor(buf, inv_dims, inv_perm, elem, n_sub, alpha)
    torch.cuda.synchronize()
    held = max(held, E.last_peak_extra_bytes())
    assert held <= alpha * nbytes, (held, nbytes)
    print("  check round trip", flush=True)
    assert E.verify_pattern(buf, dims, list(range(1, len(dims) + 1)), elem, seed, 1) == 0
    del buf
    _free(torch)
    return held


def test_high_order_6d_full_size(torch_cuda):
    # dims (32,48,64,40,36,8), perm (6 4 2 5 1 3), uint32 from seed 1, 4.53 GB
    held = _full_size(torch_cuda, [32, 48, 64, 40, 36, 8], [6, 4, 2, 5, 1, 3], 4, 1)
    print(f"  extra device bytes held: {held}", flush=True)


def test_beyond_half_hbm_172GB(torch_cuda):
    # dims (2,163840,65536,2), perm (1 3 2 4), uint32 from seed 2, 171.8 GB:
    # an out-of-place copy could not fit beside it
    held = _full_size(torch_cuda, [2, 163840, 65536, 2], [1, 3, 2, 4], 4, 2)
    print(f"  extra device bytes held: {held}", flush=True)


def test_extra_memory_follows_alpha(torch_cuda):
    # tables and arena held for a large tensor are given back when a small
    # tensor's call has a smaller budget (with no holding floor), and kept
    # within the floor otherwise
    import eithot_amd as E
    torch = torch_cuda
    _free(torch)
    floor = E.set_tuning("hold_floor", 0)
    big = [32, 48, 64, 40, 36, 8]
    vol = int(np.prod(big))
    buf = torch.empty(vol * 4, dtype=torch.uint8, device="cuda:0")
    E.fill_pattern(buf, vol, 4, 1, 1)
    E.transpose_tensor(buf, big, [6, 4, 2, 5, 1, 3], 4, 1, 0.05)
    held_big = E.last_peak_extra_bytes()
    assert 0 < held_big <= 0.05 * vol * 4
    small = [128, 8, 36, 5]  # a cycle pass with a table of its own
    svol = int(np.prod(small))
    host = O.fill(svol, 4, 3, 1)
    t = torch.from_numpy(host.copy()).to("cuda:0")
    E.transpose_tensor(t, small, [1, 3, 2, 4], 4, 1, 0.05)
    torch.cuda.synchronize()
    held_small = E.last_peak_extra_bytes()
    assert held_small <= 0.05 * svol * 4, held_small
    assert_bitexact(t.cpu().numpy(), O.permute(host, small, [0, 2, 1, 3], 4), "small tensor")
    # the large tensor's plan uploads its tables again and stays correct
    E.transpose_tensor(buf, [8, 40, 48, 36, 32, 64], [5, 3, 6, 2, 4, 1], 4, 1, 0.05)
    assert E.verify_pattern(buf, big, [1, 2, 3, 4, 5, 6], 4, 1, 1) == 0
    # with the default floor of 64 MiB the small call keeps the large
    # plan's tables
    E.set_tuning("hold_floor", 64 << 20)
    try:
        E.transpose_tensor(t, [128, 36, 8, 5], [1, 3, 2, 4], 4, 1, 0.05)
        torch.cuda.synchronize()
        assert 0.05 * svol * 4 < E.last_peak_extra_bytes() <= 64 << 20
        assert_bitexact(t.cpu().numpy(), host, "small tensor, back")
    finally:
        E.set_tuning("hold_floor", floor)
    del buf
    _free(torch)


def test_batched_4096_tensors(torch_cuda):
    # 4096 tensors from seed 3: orders 2..6, dims log-uniform [2, 512], at
    # most 4 Mi elements, 2/4/8-byte elements; every tensor checked
    import eithot_amd as E
    from eithot_amd.workloads import batch_specs
    torch = torch_cuda
    _free(torch)
    specs = batch_specs(count=4096)
    offs, total = [], 0
    for dims, _, elem in specs:
        offs.append(total)
        total += (int(np.prod(dims)) * elem + 255) // 256 * 256
    buf = torch.empty(total, dtype=torch.uint8, device="cuda:0")
    base = buf.data_ptr()
    for k, ((dims, _, elem), off) in enumerate(zip(specs, offs)):
        E.fill_pattern(base + off, int(np.prod(dims)), elem, 3000 + k, 1)
    torch.cuda.synchronize()
    b = E.Batch()
    for (dims, perm, elem), off in zip(specs, offs):
        b.add(base + off, dims, perm, elem)
    print("batch of 4096: build with autotuning", flush=True)
    b.build(alpha=0.05, tune=True)
    print("batch of 4096: run", flush=True)
    b.run()
    torch.cuda.synchronize()
    bad = [k for k, ((dims, perm, elem), off) in enumerate(zip(specs, offs))
           if E.verify_pattern(base + off, dims, perm, elem, 3000 + k, 1)]
    assert not bad, f"{len(bad)} tensors differ, first {bad[:5]}"
    b.close()
    del buf
    _free(torch)


def test_served_poisson_stream(torch_cuda):
    # Poisson arrivals at 300 requests/s from seed 4, tensors 1 MB - 4 GB,
    # orders 3..6, 1/2/4/8-byte elements, scratch pool 5 % of the resident
    # bytes: every result checked, and a sample equal to the same request
    # run alone through the plain entry point
    import time

    import eithot_amd as E
    from eithot_amd.workloads import arrivals, serve_specs
    torch = torch_cuda
    _free(torch)
    count = 96
    specs = serve_specs(count=count)
    offs, total = [], 0
    for dims, _, elem in specs:
        offs.append(total)
        total += (int(np.prod(dims)) * elem + 255) // 256 * 256
    buf = torch.empty(total, dtype=torch.uint8, device="cuda:0")
    base = buf.data_ptr()
    resident = sum(int(np.prod(d)) * e for d, _, e in specs)
    for k, ((dims, _, elem), off) in enumerate(zip(specs, offs)):
        E.fill_pattern(base + off, int(np.prod(dims)), elem, 4000 + k, 1)
    torch.cuda.synchronize()
    srv = E.Server(int(0.05 * resident), max_inflight=16, nstreams=4)
    arr = arrivals(4, 300, count)
    print(f"serving {count} requests, {resident / 1e9:.1f} GB resident", flush=True)
    t0 = time.perf_counter()
    ids = []
    for k in range(count):
        while time.perf_counter() - t0 < arr[k]:
            srv.poll(False)
        ids.append(srv.submit(base + offs[k], *specs[k]))
    srv.drain()
    for k, rid in enumerate(ids):
        st, lat = srv.status(rid)
        assert st == 2 and lat > 0, (k, st)
    srv.close()
    bad = [k for k, ((dims, perm, elem), off) in enumerate(zip(specs, offs))
           if E.verify_pattern(base + off, dims, perm, elem, 4000 + k, 1)]
    assert not bad, f"served tensors differ: {bad[:5]}"
    # alone vs in the mix: the smallest requests again, one at a time
    sample = sorted(range(count), key=lambda k: int(np.prod(specs[k][0])) * specs[k][2])[:8]
    for k in sample:
        dims, perm, elem = specs[k]
        n = int(np.prod(dims)) * elem
        t = torch.empty(n, dtype=torch.uint8, device="cuda:0")
        E.fill_pattern(t, int(np.prod(dims)), elem, 4000 + k, 1)
        E.transpose_tensor(t, dims, perm, elem, 1, 0.05)
        torch.cuda.synchronize()
        assert torch.equal(t, buf[offs[k]:offs[k] + n]), f"request {k} alone differs from the mix"
    del buf
    _free(torch)


PARTITION_CASES = [
    # (2,4096,1024,2) perm 1 3 2 4 in 4 sub-tensors (v(i) = i known answer)
    ([2, 4096, 1024, 2], [1, 3, 2, 4], 4, 4, 1.0, 0),
    # prime split dims: a ragged last sub-tensor
    ([3, 100003], [2, 1], 4, 4, 0.05, 1),
    ([100003, 3], [2, 1], 8, 5, 0.05, 1),
    ([7, 1009, 5], [3, 1, 2], 2, 3, 0.2, 1),
    ([5, 6, 331, 4], [4, 3, 1, 2], 1, 5, 0.5, 1),
]


@pytest.mark.parametrize("dims,perm,elem,nsub,alpha,mode", PARTITION_CASES)
def test_partitioned_transposes(torch_cuda, dims, perm, elem, nsub, alpha, mode):
    import eithot_amd as E
    from gpu_util import run_case
    n, d, q, r = E.plan_partition(dims, perm, elem, nsub, alpha)
    assert n == nsub and (n - 1) * q + r == dims[d], (n, d, q, r)
    got, want, peak = run_case(torch_cuda, dims, perm, elem, seed=21, mode=mode, n_sub=nsub,
                               alpha=alpha)
    assert_bitexact(got, want, f"n_sub {nsub} {dims} {perm} elem {elem}")
    assert peak <= max(alpha * got.nbytes, 1 << 20)


def test_partitioned_8GB(torch_cuda):
    # (2,16384,65536,2) perm 1 3 2 4, 8.6 GB, in 64 sub-tensors
    import eithot_amd as E
    torch = torch_cuda
    _free(torch)
    dims, perm, elem, seed = [2, 16384, 65536, 2], [1, 3, 2, 4], 4, 22
    n, _, _, _ = E.plan_partition(dims, perm, elem, 64, 0.05)
    assert n == 64
    vol = int(np.prod(dims))
    buf = torch.empty(vol * elem, dtype=torch.uint8, device="cuda:0")
    E.fill_pattern(buf, vol, elem, seed, 1)
    print("partitioned 8.6 GB: forward", flush=True)
    E.transpose_tensor(buf, dims, perm, elem, 64, 0.05)
    torch.cuda.synchronize()
    assert E.last_peak_extra_bytes() <= 0.05 * vol * elem
    assert E.verify_pattern(buf, dims, perm, elem, seed, 1) == 0
    _sampled_check(torch, buf, dims, perm, elem, seed, 1)
    del buf
    _free(torch)


@pytest.mark.parametrize("reorder", [1, 2, 3])
def test_decomposer_orders(torch_cuda, reorder):
    # pre-order, post-order and the reference's mix rule as the executed plan
    import eithot_amd as E
    from gpu_util import run_case
    old = E.set_tuning("reorder", reorder)
    try:
        for dims, perm, elem in (([32, 48, 16, 40, 9, 8], [6, 4, 2, 5, 1, 3], 4),
                                 ([8, 40, 48, 9, 32, 16], [5, 3, 6, 2, 4, 1], 4),
                                 ([2, 4096, 1024, 2], [1, 3, 2, 4], 4),
                                 ([13, 15, 10, 14, 10, 14], [4, 2, 3, 1, 5, 6], 1),
                                 ([6, 5, 7, 9, 11], [2, 4, 3, 1, 5], 8)):
            stages = E.decompose(dims, perm, reorder if reorder < 3 else 0)
            assert stages
            got, want, _ = run_case(torch_cuda, dims, perm, elem, seed=23, alpha=0.05)
            assert_bitexact(got, want, f"reorder {reorder} {dims} {perm}")
    finally:
        E.set_tuning("reorder", old)
