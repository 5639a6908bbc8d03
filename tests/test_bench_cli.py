"""bench.py's command-line contract on the CPU: more than one GPU is refused
with a message and exit status 2, and a non-positive step count is refused,
both before any device work."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    p = run("--gpus", "2")
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout == ""


def test_steps_must_be_positive():
    p = run("--steps", "0")
    assert p.returncode != 0
    assert "--steps" in p.stderr
