"""The C ABI's host code under AddressSanitizer and UndefinedBehaviorSanitizer.

librvs_san.so (rpmplv2_amd.build.build_sanitized) is the engine with its
host code (rvs_capi.cpp: argument checks, staging, graph cache, async
ordering, scratch carving; rvs_server.cpp: the server thread, request queue,
memory pool, pinned staging, motion ring) instrumented; the kernels are the
normal objects.  Two host programs that link the sanitizer runtime drive it
on the GPU:
  - space_test_san: the C++ host class over the C ABI (every query form);
    its results must equal the unsanitized build's bit for bit;
  - rvs_serve_replay_san: the request server under eight planner threads
    with moving obstacles.
A sanitizer report makes the program exit non-zero (halt_on_error).  Leak
detection is off: the HIP runtime keeps allocations until exit by design.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from rpmplv2_amd import rvs, serving, workloads as W
from test_host_space import _read_results, _write_problem

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SPACE = os.path.join(HERE, "host", "_build", "space_test")
SPACE_SAN = os.path.join(HERE, "host", "_build", "space_test_san")
REPLAY_SAN = os.path.join(ROOT, "rpmplv2_amd", "_build", "san", "rvs_serve_replay_san")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


def _clean(r):
    text = r.stdout + r.stderr
    return r.returncode == 0 and "AddressSanitizer" not in text and "runtime error" not in text


def test_space_class_under_sanitizers(tmp_path):
    assert os.path.exists(SPACE_SAN), "space_test_san is not built: run __graft_entry__.build()"
    w = W.workload2(128)
    rb, sc, q = w.robot, w.scene, w.q
    np_ref = np.zeros((sc.num_obj, rb.num_links, 6), np.float32)
    np_ref[..., 5] = 0.5
    inp = str(tmp_path / "in.bin")
    _write_problem(inp, rb, sc, rvs.ENV_AABB, q, np_ref)
    outs = []
    for exe in (SPACE, SPACE_SAN):
        out = str(tmp_path / (os.path.basename(exe) + ".bin"))
        print(f"running {os.path.basename(exe)}", flush=True)
        r = subprocess.run([exe, inp, out], capture_output=True, text=True, timeout=120, env=ENV)
        print(r.stdout[-2000:], r.stderr[-4000:])
        assert _clean(r), exe
        outs.append(_read_results(out, len(q), rb.num_links, sc.num_obj))
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k], equal_nan=True), k


def test_serving_under_sanitizers(tmp_path):
    assert os.path.exists(REPLAY_SAN), "rvs_serve_replay_san is not built: run __graft_entry__.build()"
    w = W.workload4(duration_s=0.3, rate=20000.0, pool=1 << 17)
    path = str(tmp_path / "serving.bin")
    serving.write_problem(w, path, memory_budget=1 << 30)
    print(f"running rvs_serve_replay_san on {len(w.arrival)} requests", flush=True)
    r = subprocess.run([REPLAY_SAN, path], capture_output=True, text=True, timeout=150, env=ENV)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert _clean(r)
    res = json.loads(r.stdout.strip().splitlines()[-1])
    print(res)
    assert res["requests"] == len(w.arrival) and res["failures"] == 0
