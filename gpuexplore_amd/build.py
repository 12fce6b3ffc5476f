"""Builds libgpuexplore.so and the GPUexplore CLI in-tree for gfx950.

Outputs:
  gpuexplore_amd/lib/libgpuexplore.so   the C ABI (include/gpuexplore.h)
  gpuexplore_amd/bin/GPUexplore         the command-line front end
Object files go to gpuexplore_amd/_build/.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib", "libgpuexplore.so")
CLI = os.path.join(HERE, "bin", "GPUexplore")
ARCH = os.environ.get("GX_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-parameter",
          "-I", os.path.join(ROOT, "include")]
# packed FP32 instructions stay disabled in all device code
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = ["gx_model.cpp", "gx_host.cpp", "gx_kernels.hip", "gx_batch.hip", "gx_serve.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + DEVICE + ["-c", "-o", out, os.path.join(CSRC, src)]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + COMMON + ["-x", "hip"] + DEVICE + ["-c", "-o", out, os.path.join(CSRC, src)]
    _run(cmd)
    return out


def build(verbose: bool = False) -> None:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB] + objs)
    _run([HIPCC] + COMMON + ["-o", CLI, os.path.join(CSRC, "gx_cli.cpp"),
                             "-L", os.path.dirname(LIB), "-lgpuexplore",
                             "-Wl,-rpath,$ORIGIN/../lib"])
    if verbose:
        print("built", LIB, CLI)


if __name__ == "__main__":
    build(verbose=True)
