"""The benchmark's command-line contract, on the CPU: one GPU only (any other
--gpus value exits with status 2 and says why, before touching a device), and
the options the legs are switched with."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""  # no JSON line


def test_options():
    r = subprocess.run([sys.executable, BENCH, "--help"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--skip-cpu", "--skip-surrogate",
                "--skip-growth", "--skip-host", "--skip-checkpoint", "--skip-walls"):
        assert opt in r.stdout, opt
