"""Executes a product plan on a host buffer with the oracle's transposes.

Each tile pass is a batched 2-D transpose; each run of column/row passes on
one grid must be exactly the C2R (rot?, rowg, colsh) or R2C (colshinv, rowd,
rotinv?) sequence, which the oracle executes with the reference's functors.
This pins the planner's lowering (views, bases, batch counts, sequences)
independently of the device kernels.
"""
import numpy as np

from oracle import pyoracle as O

C2R = [(2, 1), (3, 5), (2, 3)]
R2C = [(2, 4), (3, 6), (2, 2)]


def simulate(buf, ops):
    view = buf.view(np.uint8)
    i = 0
    while i < len(ops):
        o = ops[i]
        if o["kind"] == 5:  # fused block permutation: out cell o <- in cell table[o]
            U, n = o["U"], o["R"]
            nb = o["B"] * n * U
            blk = view[o["base"]:o["base"] + nb].reshape(o["B"], n, U)
            view[o["base"]:o["base"] + nb] = blk[:, o["table"], :].reshape(-1)
            i += 1
            continue
        if o["kind"] == 6:  # segmented tile: gather each matrix's segments, transpose, scatter
            B, R, C, U, sb = o["B"], o["R"], o["C"], o["U"], o["segb"]
            nseg = R * C * U // sb

            def digits_sum(count, digits):
                idx = np.arange(count, dtype=np.int64)
                out = np.zeros(count, dtype=np.int64)
                for radix, stride in digits:
                    out += (idx % radix) * stride
                    idx //= radix
                assert not idx.any(), "index beyond its digits"
                return out

            addr = (digits_sum(B, o["outer"])[:, None, None] +
                    digits_sum(nseg, o["inner"])[None, :, None] +
                    np.arange(sb, dtype=np.int64)[None, None, :]).reshape(B, -1)
            assert len(np.unique(addr)) == addr.size, "segments overlap"
            m = view[addr].reshape(B, R, C, U)
            view[addr] = m.transpose(0, 2, 1, 3).reshape(B, -1)
            i += 1
            continue
        if o["kind"] == 4 and "maps" in o:  # cycle following: its cell maps in pass order
            nb = o["B"] * o["R"] * o["C"] * o["U"]
            blk = np.ascontiguousarray(view[o["base"]:o["base"] + nb])
            cells = nb // o["U"]
            for bc, R, C, m in o["maps"]:
                assert O.transpose2d_direct(blk, cells // bc, R, C, m * o["U"]) == 0
            view[o["base"]:o["base"] + nb] = blk
            i += 1
            continue
        if o["kind"] in (1, 4):  # tile, or the same transpose by cycle following
            sub = view[o["base"]:]
            rc = O.transpose2d_direct(np.ascontiguousarray(sub), 0, 0, 0, 1)  # no-op warmup
            nb = o["B"] * o["R"] * o["C"] * o["U"]
            blk = np.ascontiguousarray(view[o["base"]:o["base"] + nb])
            assert O.transpose2d_direct(blk, o["B"], o["R"], o["C"], o["U"]) == 0
            view[o["base"]:o["base"] + nb] = blk
            i += 1
            continue
        key = (o["base"], o["B"], o["R"], o["C"], o["U"])
        grp = [o]
        j = i + 1
        while j < len(ops) and ops[j]["kind"] not in (1, 4, 5) and (
                ops[j]["base"], ops[j]["B"], ops[j]["R"], ops[j]["C"], ops[j]["U"]) == key and \
                len(grp) < 3 and (ops[j]["kind"], ops[j]["map"]) != (2, 1) and \
                (ops[j]["kind"], ops[j]["map"]) != (2, 4):
            grp.append(ops[j])
            j += 1
        seq = [(g["kind"], g["map"]) for g in grp]
        m, n, c = o["R"], o["C"], o["c"]
        nb = o["B"] * m * n * o["U"]
        blk = np.ascontiguousarray(view[o["base"]:o["base"] + nb])
        if seq == (C2R if c > 1 else C2R[1:]):
            assert O.transpose2d_catanzaro(blk, o["B"], m, n, o["U"], 0) == 0
        elif seq == (R2C if c > 1 else R2C[:2]):
            # R2C acts on the C x R view of an R x C input: R = n, C = m
            assert O.transpose2d_catanzaro(blk, o["B"], n, m, o["U"], 1) == 0
        else:
            raise AssertionError(f"unexpected pass sequence {seq}")
        view[o["base"]:o["base"] + nb] = blk
        i = j
