"""CPU tests of bench.py's command line: it measures one GPU only and says so
when asked for more, and it refuses unknown workloads, both before anything
touches a GPU or loads the library."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, cwd=ROOT, timeout=120)


def test_more_than_one_gpu_is_refused_with_a_message():
    for n in ("2", "8"):
        r = run("--gpus", n, "--steps", "1", "--warmup", "0")
        assert r.returncode == 2
        assert r.stdout == ""                       # no JSON line is printed
        assert f"--gpus {n}" in r.stderr and "one GPU" in r.stderr


def test_unknown_workload_is_refused():
    r = run("--workload", "no-such-workload")
    assert r.returncode == 2 and "invalid choice" in r.stderr and r.stdout == ""
