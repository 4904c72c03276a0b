"""The runtime's plan cache (eviction never frees a plan a caller holds) and
the serving scratch pool (whole granules, refusal of requests it can never
hold), built on the host with AddressSanitizer and UndefinedBehaviorSanitizer."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_plan_cache_and_pool_under_asan(tmp_path):
    exe = str(tmp_path / "host_structs")
    subprocess.check_call(["g++", "-O1", "-g", "-std=c++17", "-pthread",
                           "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
                           "-o", exe, os.path.join(HERE, "host_structs_harness.cpp")])
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host structs ok" in r.stdout
