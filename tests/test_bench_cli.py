"""bench.py's contract on the host: one GPU only (a request for more exits
2 with a message), and no silent run without a device (exit 1)."""

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=300, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_no_device_exits_nonzero():
    import torch  # noqa: F401  (device probe as the tests' GPU fixture does)
    if torch.cuda.is_available():
        pytest.skip("a device is visible")
    r = run("--steps", "1", "--warmup", "0")
    assert r.returncode == 1
    assert "no HIP device visible" in r.stderr
    assert r.stdout.strip() == ""
