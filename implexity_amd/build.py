"""Build the HIP library ``implexity_amd/libimplexity_lbm.so`` for gfx950.

Each precision is its own translation unit so they compile in parallel; the
double-precision unit (and the host driver) use -ffp-contract=off so that FP64 results
follow the reference's IEEE operation order.  Usage: ``python -m implexity_amd.build``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libimplexity_lbm.so")
ARCH = os.environ.get("IMPLEXITY_OFFLOAD_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
UNITS = {
    "kern_f64.hip": ["-ffp-contract=off"],
    "kern_f32.hip": [],
    "kern_f16s.hip": [],
    "kern_f16s_l16.hip": [],
    "kern_bf16s.hip": [],
    "kern_jvp.hip": ["-ffp-contract=off"],
    "api.hip": ["-ffp-contract=off"],
    "revolve.cpp": ["-x", "hip"],
}


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found: the HIP library cannot be built")
    return path


def _compile(unit: str) -> str:
    src = os.path.join(CSRC, unit)
    obj = os.path.join(BUILD, unit.rsplit(".", 1)[0] + ".o")
    cmd = [hipcc(), *COMMON, *UNITS[unit], "-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compiling {unit} failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(UNITS), os.cpu_count() or 4)) as pool:
        objs = list(pool.map(_compile, UNITS))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"linking failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
