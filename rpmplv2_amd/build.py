"""Builds librvs.so (the gfx950 engine) in-tree with hipcc.

    python -m rpmplv2_amd.build               # compile rpmplv2_amd/librvs.so
    python -m rpmplv2_amd.build --sanitized   # and the host-sanitized variant

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


# Host-code sanitizers (AddressSanitizer and UndefinedBehaviorSanitizer) for
# the C ABI and the request server: their host code is instrumented, the
# kernels are the normal objects (device code is not instrumented).  The
# library leaves the sanitizer runtime to the executables that load it, which
# link it statically: the C++ host-class test and the serving replay driver.
SAN_DIR = os.path.join(BUILD_DIR, "san")
SAN_LIB = os.path.join(SAN_DIR, "librvs_san.so")
SAN_HOST = "-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined".split()
SAN_EXE_FLAGS = ["-O1", "-g", "-std=c++17", "-fno-omit-frame-pointer", "-fsanitize=address,undefined"]
CLANGXX = os.environ.get("RVS_CLANGXX", "/opt/rocm/llvm/bin/clang++")


def build_sanitized(verbose: bool = False) -> str:
    """librvs_san.so plus the sanitized host programs
    (rpmplv2_amd/_build/san/rvs_serve_replay_san,
    tests/host/_build/space_test_san); needs build() first (kernel objects)."""
    os.makedirs(SAN_DIR, exist_ok=True)
    objs = []
    for src in SOURCES:
        if src.endswith(".cpp"):
            obj = os.path.join(SAN_DIR, os.path.splitext(src)[0] + ".o")
            cmd = [HIPCC, "-x", "hip"] + COMMON + ["-O1", "-g", "-fno-omit-frame-pointer"] + SAN_HOST + \
                  ["-c", os.path.join(CSRC, src), "-o", obj]
            if verbose:
                print(" ".join(cmd))
            _run(cmd)
        else:
            obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
    tmp = SAN_LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, SAN_LIB)
    inc = ["-I", os.path.join(ROOT, "include")]
    exe = os.path.join(SAN_DIR, "rvs_serve_replay_san")
    _run([CLANGXX] + SAN_EXE_FLAGS + inc + ["-pthread", os.path.join(CSRC, "serve_replay.cpp"), "-o", exe + ".tmp",
                                           SAN_LIB, "-Wl,-rpath," + SAN_DIR])
    os.replace(exe + ".tmp", exe)
    host_out = os.path.join(ROOT, "tests", "host", "_build")
    os.makedirs(host_out, exist_ok=True)
    exe = os.path.join(host_out, "space_test_san")
    _run([CLANGXX] + SAN_EXE_FLAGS + inc + [os.path.join(ROOT, "tests", "host", "space_test.cpp"), "-o", exe + ".tmp",
                                           SAN_LIB, "-Wl,-rpath," + SAN_DIR])
    os.replace(exe + ".tmp", exe)
    return SAN_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    if "--sanitized" in sys.argv:
        print(build_sanitized(verbose="-v" in sys.argv))
