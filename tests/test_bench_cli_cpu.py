"""bench.py's command line, CPU only: the one-GPU contract and the modes it offers."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    res = run("--gpus", "2")
    assert res.returncode == 2
    assert "one GPU" in res.stderr and "--gpus 2" in res.stderr
    assert res.stdout == ""  # no JSON line


def test_modes_are_listed():
    res = run("--help")
    assert res.returncode == 0
    for flag in ("--steps", "--warmup", "--perm", "--uneven", "--graph", "--no-autotune"):
        assert flag in res.stdout, flag
