"""The command-line drivers and the benchmark workflow on the host: the
Statistic summary against the reference's published experiment-2 figures,
and Nd_estimation (reference policy) against its 57 published outputs,
through the workflow's estimation stage."""
import json
import os
import subprocess

import pytest

from eithot_amd.cli import bench_workflow as W

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
BIN = os.path.join(HERE, "..", "eithot_amd", "bin")


def test_statistic_reproduces_published_exp2_summary():
    # mean and sample standard deviation of the 57 published Execution
    # Times, and of the throughputs 2 * data size / time (Example_png/exp2_16.png)
    runs = json.load(open(os.path.join(GOLD, "exp2_published_runs.json")))
    pub = runs["published_summary"]
    t = [c["total_ms"] for c in runs["cases"]]
    g = [2 * c["data_gb"] * 1e3 / c["total_ms"] for c in runs["cases"]]
    assert len(t) == 57
    tm, ts = W.statistic(t)
    gm, gs = W.statistic(g)
    assert abs(tm - pub["time_mean_ms"]) < 5e-4 and abs(ts - pub["time_std_ms"]) < 5e-4
    assert abs(gm - pub["throughput_mean_gbs"]) < 1e-3 and abs(gs - pub["throughput_std_gbs"]) < 1e-3
    assert W.statistic([]) != W.statistic([1.0])
    assert W.statistic([4.0]) == (4.0, 0.0)


def test_statistic_file(tmp_path):
    (tmp_path / "inplace_bench.txt").write_text("10.00000\n12.00000\n14.00000\n")
    (tmp_path / "inplace_bench_throughput.txt").write_text("5.00000\n6.00000\n7.00000\n")
    s = W.write_statistic(str(tmp_path))
    assert s["cases"] == 3 and s["time_mean_ms"] == 12.0 and s["time_std_ms"] == 2.0
    assert s["throughput_mean_gbs"] == 6.0 and s["throughput_std_gbs"] == 1.0
    txt = (tmp_path / "Statistic.txt").read_text()
    assert "Execution Time mean: 12.00000 ms, sample std: 2.00000 ms" in txt


@pytest.mark.skipif(not os.path.exists(os.path.join(BIN, "Nd_estimation")),
                    reason="drivers not built")
def test_nd_estimation_workflow_reproduces_exp2(tmp_path, capsys):
    # the 57 exp2 input lines through the workflow's estimation stage with
    # the reference policy: every printed init size and every
    # proper_input.txt line equals the published output
    gold = json.load(open(os.path.join(GOLD, "estimator_exp2.json")))
    cases = gold["cases"]
    src = tmp_path / "exp2.txt"
    src.write_text("".join(" ".join(map(str, c["dims"] + c["perm"])) + " 4 1 0.01\n"
                           for c in cases))
    env = dict(os.environ, EITHOT_ESTIMATOR="reference")
    stats = W.run_case_file(str(src), 2, str(tmp_path / "out"), estimate_only=True, env=env)
    assert stats == {"cases": 57, "failed": 0}
    out = capsys.readouterr().out
    inits = [float(line.split("=")[1].split()[0]) for line in out.splitlines()
             if line.startswith("init Extra Memory size")]
    assert inits == [c["init_mb"] for c in cases]
    assert out.count("new ratio = 0.010000") == 57
    assert "Final Extra Memory Size = 3.71098 MB, take 0.02000 percent input data size" in out
    got = (tmp_path / "out" / "exp2" / "proper_input.txt").read_text().splitlines()
    want = [" ".join(map(str, c["dims"] + c["perm"])) + f" 4 {c['nb']} {c['alpha']:.2f}"
            for c in cases]
    assert got == want


@pytest.mark.skipif(not os.path.exists(os.path.join(BIN, "Nd_estimation")),
                    reason="drivers not built")
def test_drivers_reject_bad_argv():
    r = subprocess.run([os.path.join(BIN, "Nd_estimation"), "4", "4"], capture_output=True,
                       text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([os.path.join(BIN, "test_inplace"), "4", "4", "1"], capture_output=True,
                       text=True)
    assert r.returncode == 2 and "usage" in r.stderr
