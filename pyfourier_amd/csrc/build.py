"""Build libpfhip.so (the HIP NUFFT engine) for gfx950 with hipcc, in-tree.

Usage: python pyfourier_amd/csrc/build.py [--jobs N]

Every source compiles to an object under pyfourier_amd/csrc/_build/ and the
objects link into pyfourier_amd/libpfhip.so.  Packed-FP32 instruction
selection is disabled for all device code.
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
OUT = os.path.join(PKG, "libpfhip.so")
BUILD = os.path.join(HERE, "_build")

SOURCES = ["pf_plan.hip", "pf_ops.hip", "pf_fftk.hip", "pf_toep.hip", "pf_grad.hip", "pf_cart.hip", "pf_grog.hip", "pf_dcf.hip", "pf_noalloc.hip", "pf_fft.cpp", "pf_api.cpp"]
ARCH = os.environ.get("PF_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def common_flags() -> list[str]:
    return [
        "-O3",
        "-fPIC",
        "-std=c++17",
        "-Wall",
        "-Wno-unused-function",
        "-Wno-unused-result",
        f"-I{os.path.join(ROOT, 'include')}",
    ]


def compile_one(src: str) -> str:
    path = os.path.join(HERE, src)
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), "-c", path, "-o", obj] + common_flags()
    if src == "pf_plan.hip":
        # plan arithmetic must round exactly like the reference (no FMA contraction)
        cmd += ["-ffp-contract=off"]
    if src.endswith(".hip"):
        cmd += [
            f"--offload-arch={ARCH}",
            "-Xclang",
            "-target-feature",
            "-Xclang",
            "-packed-fp32-ops",
        ]
    else:
        cmd += ["-D__HIP_PLATFORM_AMD__"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(jobs: int = 4, verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lrocfft"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args()
    try:
        build(a.jobs)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
