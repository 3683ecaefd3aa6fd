"""bench.py's command-line contract on the host: more than one GPU is
refused with a message and a nonzero status (the benchmark runs on one GPU),
and the options the measurement commands under profiles/ use exist."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True,
                          text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode != 0
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_documented_options_exist():
    r = run("--help")
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--no-cpu", "--kernel-steps", "--no-autotune", "--copy-calibration",
                "--ensemble", "--t-end", "--seq-steps", "--adjoint", "--adj-batch", "--adj-checkpoint",
                "--adj-steps", "--adj-fd", "--adj-check-steps"):
        assert opt in r.stdout, opt
