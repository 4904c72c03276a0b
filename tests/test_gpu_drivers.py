"""The command-line drivers on the GPU: test_inplace on the 64 MB 4-D
workload in both of the reference's memory modes, and the whole benchmark
workflow (Nd_estimation -> proper_input.txt -> test_inplace -> Statistic) on
a short case file."""
import os
import subprocess

import pytest

from eithot_amd.cli import bench_workflow as W

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "eithot_amd", "bin")


def _run(args, cwd, env=None):
    print("run", " ".join(args), flush=True)
    r = subprocess.run([os.path.join(BIN, "test_inplace")] + args, cwd=cwd, capture_output=True,
                       text=True, env=env, timeout=240)
    print(r.stdout, r.stderr, flush=True)
    return r


def test_test_inplace_64MB(torch_cuda, tmp_path):
    r = _run("2 4096 1024 2 1 3 2 4 4 1 0.05".split(), str(tmp_path))
    assert r.returncode == 0
    lines = r.stdout.splitlines()
    assert lines[0] == "4D Inplace 1324 transpose"
    assert lines[1] == "number of sub_tensors = 1"
    assert lines[2] == "Dims = 2 4096 1024 2, Data size = 0.06711 GB"
    assert "Inplace transpose verify_inplace" in lines and "Correct transpose" in lines
    assert any(x.startswith("Execution Time without h2d2h: ") for x in lines)
    assert any(x.startswith("Execution Time: ") for x in lines)
    for log in ("inplace_bench.txt", "inplace_bench_throughput.txt"):
        vals = (tmp_path / log).read_text().split()
        assert len(vals) == 1 and float(vals[0]) > 0


def test_test_inplace_managed_memory(torch_cuda, tmp_path):
    env = dict(os.environ, EITHOT_MEMORY_TYPE="managed")
    r = _run("2 512 256 2 1 3 2 4 4 1 0.05".split(), str(tmp_path), env)
    assert r.returncode == 0
    assert r.stdout.startswith("4D Inplace 1324 transpose, Data size = ")
    assert "Correct transpose" in r.stdout.splitlines()


def test_bench_workflow_end_to_end(torch_cuda, tmp_path):
    cases = tmp_path / "exp_small.txt"
    cases.write_text("7248 7248 2 1 4 1 0.01\n"
                     "96 75 96 75 3 2 4 1 4 1 0.01\n"
                     "32 15 15 32 15 15 4 3 1 6 2 5 4 1 0.01\n")
    stats = W.run_case_file(str(cases), 2, str(tmp_path / "out"))
    assert stats["cases"] == 3 and stats["failed"] == 0
    save = tmp_path / "out" / "exp_small"
    for f in ("inplace_bench.txt", "inplace_bench_throughput.txt", "Statistic.txt",
              "proper_input.txt"):
        assert (save / f).exists(), f
    assert len((save / "proper_input.txt").read_text().splitlines()) == 3
    assert "cases: 3" in (save / "Statistic.txt").read_text()
