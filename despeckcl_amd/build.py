"""Build libdespeckcl_amd.so in-tree with hipcc for gfx950.

    python -m despeckcl_amd.build        (or despeckcl_amd.build.build())

Device code is compiled without FP contraction (explicit fmaf() marks every fused
multiply-add) and without packed FP32 instructions.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdespeckcl_amd.so")
BUILD_DIR = os.path.join(ROOT, "_build", "despeckcl_amd")

ARCH = os.environ.get("DKX_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

SOURCES = ["nlsar.hip", "nlinsar.hip", "server.hip", "noalloc.hip", "capi.cpp", "batch.cpp", "nlinsar_batch.cpp", "nlinsar_simu.cpp", "tune.cpp", "stats.cpp", "stats_io.cpp", "training.hip", "boxcar.hip", "goldstein.hip",
           "despeckcl_api.cpp"] + [
    f"nlsar_pass_{kind}_{shape}.hip" for kind in ("stat", "last")
    for shape in ("t32r4_w10", "t32r2_w10", "t32r1_w10", "t32r4_w20", "t32r2_w10_cov2", "t32r4_w20_cov2",
                  "t32r2_w10_cov3", "t32r4_w20_cov3")]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
DEVICE = ["-x", "hip", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-fgpu-rdc" if False else "-fno-gpu-rdc"]


def _compile(src):
    obj = os.path.join(BUILD_DIR, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON
    if src.endswith(".hip"):
        cmd += DEVICE
    else:
        cmd += ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(BUILD_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
