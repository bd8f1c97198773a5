"""The benchmark program's command line, on the CPU: it runs on one GPU and
refuses more with a message (before touching a device), and its options parse."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    for n in ("2", "8"):
        r = run_bench("--gpus", n)
        assert r.returncode == 2
        assert f"--gpus {n} is not supported" in r.stderr
        assert r.stdout == ""


def test_help_lists_the_contract_options():
    r = run_bench("--help")
    assert r.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--skeleton-reps"):
        assert opt in r.stdout
