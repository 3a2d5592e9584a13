"""Builds librvs.so (the gfx950 engine) in-tree with hipcc.

    python -m rpmplv2_amd.build          # compile rpmplv2_amd/librvs.so

Compile flags:
  --offload-arch=gfx950                    MI355X only
  -ffp-contract=off                        no FMA contraction, so the exact
                                           kernels keep the reference's rounding
                                           (those sources also carry '#pragma
                                           clang fp contract(off)'); only the
                                           FP32 bulk pass (rvs_fast.hip) is
                                           built with -ffp-contract=fast
  -target-feature -packed-fp32-ops         no packed FP32 instructions
Division and square root stay IEEE (hipcc's default
-fhip-fp32-correctly-rounded-divide-sqrt), denormals are kept.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librvs.so")
BUILD_DIR = os.path.join(HERE, "_build")

SOURCES = ["rvs_kernels.hip", "rvs_edges.hip", "rvs_fast.hip", "rvs_spines.hip", "rvs_traj.hip", "rvs_capi.cpp", "rvs_server.cpp"]
# the FP32 bulk pass is the one translation unit that allows FMA contraction
EXTRA = {"rvs_fast.hip": ["-ffp-contract=fast"]}
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
COMMON = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(ROOT, "include"), "-Wall"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + COMMON + EXTRA.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd[1:1] = ["-x", "hip"]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    # the serving replay driver (a host program over the C ABI)
    exe = os.path.join(BUILD_DIR, "rvs_serve_replay")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-pthread", "-I", os.path.join(ROOT, "include"),
          os.path.join(CSRC, "serve_replay.cpp"), "-o", exe + ".tmp", LIB, "-Wl,-rpath," + HERE])
    os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
