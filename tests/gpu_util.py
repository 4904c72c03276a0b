"""Helpers of the GPU parity tests: device round trips checked by the oracle."""
import numpy as np

from oracle import pyoracle as O

NP = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


def run_case(torch, dims, perm1, elem, seed=1, mode=1, n_sub=1, alpha=0.05):
    """Transpose a seeded tensor on the GPU; return (got, want, peak_extra)."""
    import eithot_amd as E
    vol = int(np.prod(dims))
    host = O.fill(vol, elem, seed, mode)
    dev = torch.from_numpy(host.copy()).to("cuda:0")
    print(f"case dims={dims} perm={perm1} elem={elem} n_sub={n_sub} alpha={alpha}", flush=True)
    E.transpose_tensor(dev, dims, perm1, elem, n_sub, alpha)
    torch.cuda.synchronize()
    got = dev.cpu().numpy()
    want = O.permute(host, dims, [p - 1 for p in perm1], elem)
    return got, want, E.last_peak_extra_bytes()


def assert_bitexact(got, want, what=""):
    if not np.array_equal(got, want):
        bad = np.flatnonzero(got != want)
        raise AssertionError(f"{what}: {bad.size} bytes differ, first at byte {bad[0]}")
