"""The GPUexplore command line: flag handling, messages and exit codes of the
reference's main() (cu:1785-1883, 1984-1986, 2000-2003)."""

import os
import subprocess

import pytest

from conftest import GOLDEN


def run(gx, *args):
    return subprocess.run([gx.CLI_PATH, *args], capture_output=True, text=True, timeout=120)


def test_no_args(gx):
    r = run(gx)
    assert r.returncode == 1
    assert "ERROR: No input network given!" in r.stderr
    assert r.stdout.startswith("Usage: GPUexplore <model> [OPTIONS]")


@pytest.mark.parametrize("flag", ["-h", "--help", "-?"])
def test_help(gx, flag):
    r = run(gx, flag)
    assert r.returncode == 0 and "Usage: GPUexplore" in r.stdout
    r = run(gx, os.path.join(GOLDEN, "tiny2"), "-d", flag)
    assert r.returncode == 0 and "Usage: GPUexplore" in r.stdout


def test_unknown_option(gx):
    r = run(gx, os.path.join(GOLDEN, "tiny2"), "--frobnicate")
    assert r.returncode == 1
    assert "ERROR: unrecognized option --frobnicate" in r.stderr


def test_missing_model(gx, model_dir):
    r = run(gx, os.path.join(model_dir, "nope"), "-d")
    assert r.returncode == 1
    assert "ERROR: input network does not exist!" in r.stderr


def test_liveness_line_with_p(gx):
    r = run(gx, os.path.join(GOLDEN, "tiny2"), "-p")
    assert "Property to check is not a liveness property" in r.stdout


def test_no_device(gx):
    if gx.device_count() > 0:
        pytest.skip("a device is visible")
    r = run(gx, os.path.join(GOLDEN, "tiny2"), "-d")
    assert r.returncode == 1
    assert "No HIP compatible GPU detected" in r.stderr


def test_unknown_por_method(gx, model_dir):
    from gpuexplore_amd import gen
    p = os.path.join(model_dir, "m")
    gen.config1(1).write(p)
    r = subprocess.run([gx.CLI_PATH, p, "--por-method", "sleep-sets"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "unknown partial-order reduction method sleep-sets" in r.stderr
