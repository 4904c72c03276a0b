"""The C-ABI library loads on a machine without a GPU and exports every
symbol include/eithot.h declares (no compute calls here)."""
import ctypes
import os
import re
import subprocess

import pytest

import eithot_amd as E
from eithot_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "eithot.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(eithot_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_python_export_lists_agree():
    assert declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = E.lib()
    for name in declared():
        assert hasattr(L, name), name
    out = subprocess.check_output(["nm", "-D", "--defined-only", _lib.LIB_PATH], text=True)
    for name in declared():
        assert re.search(rf"\bT {name}\b", out), name


def test_library_is_a_gfx950_hip_build():
    # the HIP fat binary of this build is embedded and targets gfx950
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data
    assert b"hipFatBinSegment" in data or b".hip_fatbin" in data


def test_error_message_thread_local_and_planning_without_gpu():
    n, a, x = E.estimate([2, 4096, 1024, 2], [1, 3, 2, 4], 4, 0.05)
    assert n == 1 and x <= 0.05 * 2 * 4096 * 1024 * 2 * 4 and a <= 0.05
    rc = E.lib().eithot_transpose(None, E.i64([2]), E.i32([1]), 1, 4, 1, 0.05)
    assert rc == -1
    assert b"null" in E.lib().eithot_last_error()


def test_tuning_rejects_out_of_range_values():
    import eithot_amd as E
    for key, bad in (("segtile_nt", 1), ("regroup", 2), ("lds_threads_auto", 5),
                     ("cycle_variant", 8), ("threads", 96), ("cycle_lmax", 1),
                     ("cycle_u_depth", 5), ("cycle_u_depth", 16), ("block_pad", 2),
                     ("min_groups", 0)):
        with pytest.raises(E.EithotError):
            E.set_tuning(key, bad)
    for key, good in (("segtile_nt", 2), ("regroup", 0), ("cycle_lmax", 64), ("cycle_u_depth", 8),
                      ("block_pad", 0), ("min_groups", 256)):
        old = E.set_tuning(key, good)
        try:
            assert E.get_tuning(key) == good
        finally:
            E.set_tuning(key, old)
    with pytest.raises(E.EithotError):
        E.set_tuning("no_such_knob", 1)
