"""The C++ host class rvs::RealVectorSpaceGPU (include/rvs_space.hpp): the
reference's single-state RealVectorSpace methods over the C ABI.

tests/host/space_test (built by __graft_entry__.build()) drives the class on
a problem written here and writes its results; they must equal the batched C
ABI (bit for bit, same arithmetic) and the oracle's verdicts.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from rpmplv2_amd import robots as R, rvs, workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BIN = os.path.join(HERE, "host", "_build", "space_test")


def test_space_header_compiles():
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-I", os.path.join(ROOT, "include"), "-x", "c++", "-"],
                       input='#include "rvs_space.hpp"\nint main() { return 0; }\n', capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def _write_problem(path, rb, scene, semantics, q, np_ref):
    desc, keep = rvs.robot_desc(rb)
    boxes = rvs.boxes_of(scene)
    q = np.ascontiguousarray(q, np.float32)
    with open(path, "wb") as f:
        f.write(np.array([rb.skeleton_kind, len(rb.joints), rb.num_links, rb.ground_included, len(rb.self_pairs),
                          scene.num_obj, semantics, rb.num_dofs], np.int32).tobytes())
        f.write(np.float32(rb.gripper_length).tobytes())
        f.write(np.int64(len(q)).tobytes())
        f.write(C.string_at(C.addressof(keep[0]), C.sizeof(keep[0])))
        f.write(np.ascontiguousarray(rb.radius, np.float32).tobytes())
        f.write(np.ascontiguousarray(np.asarray(rb.self_pairs, np.int32).reshape(-1)).tobytes())
        f.write(C.string_at(C.addressof(boxes), C.sizeof(boxes)))
        f.write(q.tobytes())
        f.write(np.ascontiguousarray(np_ref, np.float32).tobytes())


def _read_results(path, n, L, M):
    b = open(path, "rb").read()
    off = 0

    def take(dt, count):
        nonlocal off
        a = np.frombuffer(b, dt, count, off)
        off += a.nbytes
        return a

    out = dict(valid=take(np.uint8, n), d=take(np.float32, n), prof=take(np.float32, n * L).reshape(n, L),
               has_np=take(np.uint8, n), under=take(np.float32, n), under_stored=take(np.uint8, n),
               self_col=take(np.uint8, n), edge=take(np.uint8, n), np0=take(np.float32, M * L * 6))
    assert off == len(b)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["xarm6", "chain22_obb"])
def test_real_vector_space_gpu_class(oracle, engine, tmp_path, case):
    assert os.path.exists(BIN), "tests/host/_build/space_test is not built: run __graft_entry__.build()"
    if case == "xarm6":
        w = W.workload2(256)
        sem, oriented = rvs.ENV_AABB, False
    else:
        w = W.workload5(64)
        sem, oriented = rvs.ENV_OBB, True
    rb, sc, q = w.robot, w.scene, w.q
    n, L, M = len(q), rb.num_links, sc.num_obj
    osc = oracle.scene_of(sc, oriented=oriented)
    od, oprof, over, onp = oracle.distance(rb, osc, q, 32, want_np=True)
    k = int(np.argmax(over == 1))
    np_ref = onp[k]
    inp, outp = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_problem(inp, rb, sc, sem, q, np_ref)
    print(f"space_test {case}: n={n}", flush=True)
    r = subprocess.run([BIN, inp, outp], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    res = _read_results(outp, n, L, M)
    rid = engine.load_robot(rb)
    eid = engine.create_env(sc, sem)
    d, prof, ver, npo = engine.distance(rid, eid, q, precision=rvs.PREC_REF32, want_nearest=True)
    assert np.array_equal(res["d"], d) and np.array_equal(res["prof"], prof)
    assert np.array_equal(res["has_np"], ver)
    assert np.array_equal(res["valid"], engine.is_valid(rid, eid, q))
    assert np.array_equal(res["valid"], oracle.is_valid(rb, osc, q))
    assert np.array_equal((res["d"] > 0).astype(np.uint8), over)
    if ver[0]:
        assert np.array_equal(res["np0"], npo[0].reshape(-1), equal_nan=True)
    ud, uprof, ust = engine.underestimation(rid, eid, q, np_ref)
    assert np.array_equal(res["under_stored"], ust)
    assert np.array_equal(res["under"], np.where(ust == 1, ud, 0).astype(np.float32))
    _, free = engine.self_collision(rid, q)
    assert np.array_equal(res["self_col"], 1 - free)
    q2 = np.roll(q, -1, axis=0)
    assert np.array_equal(res["edge"], engine.edge_is_valid(rid, eid, q, q2))
