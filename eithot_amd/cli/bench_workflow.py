"""The reference's benchmark workflow over this build's command-line drivers.

    python -m eithot_amd.cli.bench_workflow CASES [CASES ...] [--exp K] [--out DIR]

Follows Benchmark/Bench.sh:25-84 step by step, one case file at a time:

* experiment 1 (``--exp 1``): every line of the case file
  (``d1..dn p1..pn elem n_sub alpha``) runs through ``test_inplace`` as given;
* every other experiment: every line first runs through ``Nd_estimation``,
  which appends the planned ``dims perm elem nb alpha`` line to
  ``proper_input.txt``, and then every line of ``proper_input.txt`` runs
  through ``test_inplace``;
* ``Statistic`` summarises the two logs ``test_inplace`` appends to
  (``inplace_bench.txt``: milliseconds, ``inplace_bench_throughput.txt``:
  GB/s) as mean and sample standard deviation in ``Statistic.txt``; the
  reference's README publishes these figures for experiment 2
  (Example_png/exp2_16.png), and the script that computes them is not in the
  reference tree, so the format here is this build's own;
* the logs and the summary move into a directory named after the case file
  (its path without ``.txt``), under ``--out``.

Each case runs in its own process, as in the reference, and its stdout is
echoed.  The binaries are the ones ``__graft_entry__.build()`` puts in
``eithot_amd/bin``.
"""
import argparse
import math
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(os.path.dirname(HERE), "bin")
LOGS = ("inplace_bench.txt", "inplace_bench_throughput.txt")


def statistic(values):
    """Mean and sample standard deviation (n - 1 in the denominator)."""
    n = len(values)
    if n == 0:
        return float("nan"), float("nan")
    mean = sum(values) / n
    if n == 1:
        return mean, 0.0
    var = sum((v - mean) ** 2 for v in values) / (n - 1)
    return mean, math.sqrt(var)


def read_log(path):
    if not os.path.exists(path):
        return []
    with open(path) as f:
        return [float(x) for x in f.read().split()]


def write_statistic(workdir):
    """Statistic.txt from the two logs in workdir; returns the figures."""
    t = read_log(os.path.join(workdir, LOGS[0]))
    g = read_log(os.path.join(workdir, LOGS[1]))
    tm, ts = statistic(t)
    gm, gs = statistic(g)
    with open(os.path.join(workdir, "Statistic.txt"), "w") as f:
        f.write(f"cases: {len(t)}\n")
        f.write(f"Execution Time mean: {tm:.5f} ms, sample std: {ts:.5f} ms\n")
        f.write(f"Throughput mean: {gm:.5f} GB/s, sample std: {gs:.5f} GB/s\n")
    return {"cases": len(t), "time_mean_ms": tm, "time_std_ms": ts, "throughput_mean_gbs": gm,
            "throughput_std_gbs": gs}


def _run(exe, line, workdir, env=None):
    args = [os.path.join(BIN, exe)] + line.split()
    r = subprocess.run(args, cwd=workdir, capture_output=True, text=True, env=env)
    sys.stdout.write(r.stdout)
    sys.stdout.flush()
    if r.returncode:
        sys.stderr.write(r.stderr)
    return r


def run_case_file(path, exp, out, estimate_only=False, env=None):
    """Runs one case file through the workflow; returns its Statistic figures."""
    name = os.path.splitext(os.path.basename(path))[0]
    save = os.path.join(out, name)
    if os.path.isdir(save):
        shutil.rmtree(save)
    os.makedirs(save)
    work = os.path.join(out, "_run")
    os.makedirs(work, exist_ok=True)
    for f in LOGS + ("proper_input.txt", "Statistic.txt"):
        p = os.path.join(work, f)
        if os.path.exists(p):
            os.remove(p)
    print(path, flush=True)
    with open(path) as f:
        lines = [x.strip() for x in f if x.strip()]
    failed = 0
    if exp != 1:
        for n, line in enumerate(lines, 1):
            print(f"Case {n}", flush=True)
            failed += _run("Nd_estimation", line, work, env).returncode != 0
            print("", flush=True)
        if estimate_only:
            shutil.move(os.path.join(work, "proper_input.txt"), save)
            return {"cases": len(lines), "failed": failed}
        with open(os.path.join(work, "proper_input.txt")) as f:
            lines = [x.strip() for x in f if x.strip()]
    for n, line in enumerate(lines, 1):
        print(f"Case {n}", flush=True)
        failed += _run("test_inplace", line, work, env).returncode != 0
        print("", flush=True)
    stats = write_statistic(work)
    stats["failed"] = failed
    for f in LOGS + ("Statistic.txt", "proper_input.txt"):
        p = os.path.join(work, f)
        if os.path.exists(p):
            shutil.move(p, save)
    return stats


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("cases", nargs="+", help="case files, one 'dims perm elem n_sub alpha' per line")
    ap.add_argument("--exp", type=int, default=2, help="experiment number (1: no estimation step)")
    ap.add_argument("--out", default=".", help="directory the per-file results go under")
    ap.add_argument("--estimate-only", action="store_true",
                    help="stop after Nd_estimation (no GPU needed)")
    args = ap.parse_args(argv)
    rc = 0
    for path in args.cases:
        stats = run_case_file(path, args.exp, args.out, args.estimate_only)
        print(stats, flush=True)
        rc |= 1 if stats.get("failed") else 0
    return rc


if __name__ == "__main__":
    sys.exit(main())
