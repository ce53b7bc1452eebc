"""bench.py command line: the single-GPU contract and the flags the measurement
legs are switched by (no GPU needed: only argument handling is exercised)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT, capture_output=True,
                          text=True, timeout=120)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "exactly one GPU" in r.stderr
    assert r.stdout == ""   # no JSON line for a run that was not made


def test_help_lists_the_measurement_legs():
    r = _run("--help")
    assert r.returncode == 0
    for flag in ("--steps", "--warmup", "--gpus", "--no-cpu", "--no-ref", "--no-library", "--no-rows", "--no-direct",
                 "--no-host", "--grad-only", "--fft-planes", "--precisions", "--method"):
        assert flag in r.stdout, flag
