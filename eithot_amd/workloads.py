"""Synthetic workloads of the benchmark configurations (seeded samplers)."""
import numpy as np


def batch_specs(seed=3, count=4096, cap=4 * 1024 * 1024):
    """Batched workload: order U{2..6}, dims log-uniform on [2, 512] with the
    tensor resampled until it holds at most `cap` elements, element size
    2/4/8 bytes with p = .25/.5/.25, permutation uniform over non-identity."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        r = int(rng.integers(2, 7))
        while True:
            dims = [int(round(np.exp(x))) for x in rng.uniform(np.log(2), np.log(512), size=r)]
            if int(np.prod(dims)) <= cap:
                break
        elem = int(rng.choice([2, 4, 8], p=[0.25, 0.5, 0.25]))
        perm = list(range(1, r + 1))
        while perm == sorted(perm):
            perm = [int(p) + 1 for p in rng.permutation(r)]
        out.append((dims, perm, elem))
    return out


def serve_specs(seed=4, count=128, lo=1 << 20, hi=4 << 30):
    """Served workload: tensor bytes log-uniform on [1 MB, 4 GB], order
    U{3..6}, element size 1/2/4/8, non-identity permutations; dims split the
    element count with random log-proportions (each dim >= 2)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nbytes = float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        elem = int(rng.choice([1, 2, 4, 8]))
        r = int(rng.integers(3, 7))
        vol = max(2 ** r, int(nbytes // elem))
        w = rng.dirichlet(np.ones(r))
        dims = [max(2, int(round(np.exp(np.log(vol) * x)))) for x in w]
        perm = list(range(1, r + 1))
        while perm == sorted(perm):
            perm = [int(p) + 1 for p in rng.permutation(r)]
        out.append((dims, perm, elem))
    return out


def arrivals(seed, rate, count):
    """Poisson arrival times (seconds from start) at `rate` requests/s."""
    rng = np.random.default_rng(seed * 1000 + int(rate))
    return np.cumsum(rng.exponential(1.0 / rate, size=count))
