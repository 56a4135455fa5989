"""bench.py's contract on a machine without a device: more than one GPU is refused with a
message, and a run finds the missing device and fails rather than measuring anything else."""
import os
import subprocess
import sys

import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT, capture_output=True,
                          text=True, timeout=300)


def test_more_than_one_gpu_is_refused_with_a_message():
    out = run("--gpus", "2")
    assert out.returncode == 2
    assert "one GPU" in out.stdout + out.stderr
    assert not out.stdout.strip().startswith("{")


@pytest.mark.skipif(torch.cuda.is_available(), reason="the device is there")
def test_a_run_without_a_device_fails_loudly():
    out = run("--steps", "1", "--warmup", "0", "--no-cpu")
    assert out.returncode != 0
    assert "no HIP device" in out.stderr
    assert "{" not in out.stdout
