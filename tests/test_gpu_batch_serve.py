"""GPU parity of the batched and served paths: every tensor equals the
oracle's permutation, alone or in the mix."""
import numpy as np
import pytest

from gpu_util import assert_bitexact
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def random_tensor(rng, orders=(2, 6), dmax=64, vmax=200_000, elems=(1, 2, 4, 8)):
    while True:
        r = int(rng.integers(orders[0], orders[1] + 1))
        dims = [int(round(np.exp(x))) for x in rng.uniform(np.log(2), np.log(dmax), size=r)]
        if np.prod(dims) <= vmax:
            break
    perm = list(range(1, r + 1))
    while perm == sorted(perm):
        perm = [int(p) + 1 for p in rng.permutation(r)]
    return dims, perm, int(rng.choice(elems))


def carve(torch, specs, align=256):
    offs, total = [], 0
    for dims, _, elem in specs:
        offs.append(total)
        total += (int(np.prod(dims)) * elem + align - 1) // align * align
    return torch.zeros(max(total, 1), dtype=torch.uint8, device="cuda:0"), offs


def test_batch_matches_oracle(torch_cuda):
    import eithot_amd as E
    torch = torch_cuda
    rng = np.random.default_rng(3)
    specs = [random_tensor(rng) for _ in range(96)]
    # tensors whose plans hold cycle passes (whole and cut cycles, composed
    # maps) and fused block passes, which run inside the batched launches
    specs += [([128, 8, 36, 5], [1, 3, 2, 4], 4), ([1536, 64, 12], [1, 3, 2], 4),
              ([48, 40, 8, 30], [3, 2, 1, 4], 4), ([12, 2, 25, 14, 53, 3], [3, 6, 4, 1, 5, 2], 8),
              ([2, 42, 2, 11, 11], [5, 2, 1, 4, 3], 8), ([32, 40, 48, 24], [1, 3, 2, 4], 4),
              # segmented tiles, tile last and tile first
              ([12, 12, 12, 16, 20], [1, 5, 3, 2, 4], 4), ([2, 32, 4, 5, 32], [5, 2, 4, 1, 3], 4)]
    kinds = set()
    for dims, perm, elem in specs:
        kinds |= {o["kind"] for o in E.plan_describe(dims, perm, elem, 1, 0.5)[0]}
    assert {4, 5, 6} <= kinds, kinds
    buf, offs = carve(torch, specs)
    hosts = []
    for k, ((dims, perm, elem), off) in enumerate(zip(specs, offs)):
        h = O.fill(int(np.prod(dims)), elem, 100 + k, 1)
        hosts.append(h)
        buf[off:off + h.size] = torch.from_numpy(h).to("cuda:0")
    torch.cuda.synchronize()
    b = E.Batch()
    base = buf.data_ptr()
    for (dims, perm, elem), off in zip(specs, offs):
        b.add(base + off, dims, perm, elem)
    print("batch of", len(specs), "tensors: build with autotuning", flush=True)
    b.build(alpha=0.5, tune=True)  # tuning runs forward + inverse, leaving the data unchanged
    for h, off in zip(hosts, offs):
        assert_bitexact(buf[off:off + h.size].cpu().numpy(), h, "tuning must restore the batch")
    info = b.info()
    assert info["graph"] and info["launches"] >= 1
    b.run()
    torch.cuda.synchronize()
    for k, ((dims, perm, elem), off, h) in enumerate(zip(specs, offs, hosts)):
        want = O.permute(h, dims, [p - 1 for p in perm], elem)
        assert_bitexact(buf[off:off + h.size].cpu().numpy(), want, f"tensor {k} {dims} {perm}")
    b.close()


def test_server_alone_and_in_the_mix(torch_cuda):
    import eithot_amd as E
    torch = torch_cuda
    rng = np.random.default_rng(4)
    specs = [random_tensor(rng, orders=(3, 6), dmax=200, vmax=2_000_000) for _ in range(40)]
    # plans with a segmented tile (a composed cycle pass around it), tile
    # last and tile first, and a fused block pass
    specs += [([12, 12, 12, 16, 20], [1, 5, 3, 2, 4], 4), ([2, 32, 4, 5, 32], [5, 2, 4, 1, 3], 4),
              ([48, 40, 8, 30], [3, 2, 1, 4], 4)]
    kinds = set()
    for dims, perm, elem in specs[-3:]:
        kinds |= {o["kind"] for o in E.plan_describe(dims, perm, elem, 1, 0.05)[0]}
    assert 6 in kinds, kinds
    # alone: one request at a time through the plain entry point
    alone = []
    for k, (dims, perm, elem) in enumerate(specs):
        h = O.fill(int(np.prod(dims)), elem, 500 + k, 1)
        t = torch.from_numpy(h).to("cuda:0")
        E.transpose_tensor(t, dims, perm, elem, 1, 0.05)
        alone.append(t.cpu().numpy())
    # in the mix: all submitted at once to a server with a small scratch pool
    buf, offs = carve(torch, specs)
    for k, ((dims, perm, elem), off) in enumerate(zip(specs, offs)):
        h = O.fill(int(np.prod(dims)), elem, 500 + k, 1)
        buf[off:off + h.size] = torch.from_numpy(h).to("cuda:0")
    torch.cuda.synchronize()
    resident = sum(int(np.prod(d)) * e for d, _, e in specs)
    srv = E.Server(scratch_pool_bytes=max(int(0.05 * resident), 1 << 20), max_inflight=8,
                   nstreams=4)
    ids = [srv.submit(buf.data_ptr() + off, dims, perm, elem)
           for (dims, perm, elem), off in zip(specs, offs)]
    srv.drain()
    for k, (rid, (dims, perm, elem), off) in enumerate(zip(ids, specs, offs)):
        st, lat = srv.status(rid)
        assert st == 2 and lat > 0, (k, st, lat)
        n = int(np.prod(dims)) * elem
        got = buf[off:off + n].cpu().numpy()
        h = O.fill(int(np.prod(dims)), elem, 500 + k, 1)
        want = O.permute(h, dims, [p - 1 for p in perm], elem)
        assert_bitexact(got, want, f"served {k}")
        assert_bitexact(got, alone[k], f"served {k} vs alone")
    srv.close()


def test_server_refuses_requests_it_cannot_admit(torch_cuda):
    import eithot_amd as E
    torch = torch_cuda
    # rows of 200003 8-byte units exceed LDS even as 1-byte slices, so the row
    # pass is staged through 200003 bytes of scratch, more than a 1 KiB pool
    # holds: the request is refused at submit, nothing runs
    t = torch.zeros(3 * 200003 * 8, dtype=torch.uint8, device="cuda:0")
    srv = E.Server(scratch_pool_bytes=1024, max_inflight=2, nstreams=1)
    with pytest.raises(E.EithotError):
        srv.submit(t.data_ptr(), [3, 200003], [2, 1], 8)
    srv.close()


def test_server_orders_requests_on_one_tensor(torch_cuda):
    # forward, then inverse, then forward again on one resident tensor, with
    # unrelated requests between them on a 4-stream server: the requests on
    # the shared bytes run in submission order
    import eithot_amd as E
    torch = torch_cuda
    dims, perm, elem = [32, 48, 16, 40, 9, 8], [6, 4, 2, 5, 1, 3], 4
    inv_dims = [dims[p - 1] for p in perm]
    inv = [0] * len(perm)
    for k, p in enumerate(perm):
        inv[p - 1] = k + 1
    vol = int(np.prod(dims))
    host = O.fill(vol, elem, 31, 1)
    t = torch.from_numpy(host.copy()).to("cuda:0")
    rng = np.random.default_rng(32)
    others = [random_tensor(rng, orders=(3, 5), dmax=100, vmax=500_000) for _ in range(6)]
    obuf, ooffs = carve(torch, others)
    ohost = []
    for k, ((d, _, e), off) in enumerate(zip(others, ooffs)):
        h = O.fill(int(np.prod(d)), e, 600 + k, 1)
        ohost.append(h)
        obuf[off:off + h.size] = torch.from_numpy(h).to("cuda:0")
    torch.cuda.synchronize()
    srv = E.Server(scratch_pool_bytes=64 << 20, max_inflight=8, nstreams=4)
    p = t.data_ptr()
    seq = [(p, dims, perm), (p, inv_dims, inv), (p, dims, perm)]
    ids = []
    for k in range(len(seq)):
        ptr, d, q = seq[k]
        ids.append(srv.submit(ptr, d, q, elem))
        for j in range(2 * k, 2 * k + 2):
            d2, q2, e2 = others[j]
            srv.submit(obuf.data_ptr() + ooffs[j], d2, q2, e2)
    srv.drain()
    for rid in ids:
        assert srv.status(rid)[0] == 2
    srv.close()
    want = O.permute(host, dims, [q - 1 for q in perm], elem)
    assert_bitexact(t.cpu().numpy(), want, "forward, inverse, forward on one tensor")
    for k, ((d, q, e), off, h) in enumerate(zip(others, ooffs, ohost)):
        assert_bitexact(obuf[off:off + h.size].cpu().numpy(),
                        O.permute(h, d, [x - 1 for x in q], e), f"other request {k}")
