"""Autotuner of the FP32 bulk pass: every launch shape gives the same
per-configuration results (the shape only changes which lane evaluates a
configuration), and rvs_autotune keeps one of the candidate shapes."""
import numpy as np
import pytest

from rpmplv2_amd import robots as R, rvs, workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("wl", ["workload2", "workload5_obb"])
def test_launch_shapes_agree(engine, wl):
    if wl == "workload2":
        w = W.workload2(1 << 16)
        sem = rvs.ENV_AABB
    else:
        w = W.workload5(4096)
        sem = rvs.ENV_OBB
    rid = engine.load_robot(w.robot)
    eid = engine.create_env(w.scene, sem)
    ref = None
    for variant in (0, 1):
        for block in (64, 128, 256):
            for per_cu in ((0,) if variant == 0 else (0, 1)):
                engine.set_tuning(rid, eid, variant, block, per_cu)
                d, prof, ver, _ = engine.distance(rid, eid, w.q, precision=rvs.PREC_FP32_REFINE)
                if ref is None:
                    ref = (d, prof, ver)
                else:
                    assert np.array_equal(ver, ref[2]), (variant, block, per_cu)
                    assert np.array_equal(d, ref[0]) and np.array_equal(prof, ref[1]), (variant, block, per_cu)
    best = engine.autotune(rid, eid, w.q)
    print(wl, "autotune:", best)
    assert best["variant"] in (0, 1) and best["block"] in (64, 128, 256) and best["ms"] > 0
    assert engine.get_tuning(rid, eid)["variant"] == best["variant"]
    d, prof, ver, _ = engine.distance(rid, eid, w.q, precision=rvs.PREC_FP32_REFINE)
    assert np.array_equal(ver, ref[2]) and np.array_equal(d, ref[0])


def test_edge_launch_shapes_agree(oracle, engine):
    """Every (waves per CU, slots per lane) shape of the persistent
    first-collision launch gives the oracle's first colliding state."""
    rng = np.random.default_rng(12)
    sc = W._boxes(rng, 8, 0.05, 0.1, 1.5)
    eid = engine.create_env(sc)
    batches = []
    for n in (10, 14):
        rb = R.chain(rng.uniform(0.05, 0.1, n), radius=0.01, self_collision=True)
        m = 300
        q1 = rng.uniform(-0.3, 0.3, (m, n)).astype(np.float32)
        u = rng.normal(size=(m, n))
        u = (u / np.linalg.norm(u, axis=1, keepdims=True)).astype(np.float32)
        batches.append(W.EdgeBatch(rb, q1, u, rng.integers(1, 200, m).astype(np.int32)))
    osc = oracle.scene_of(sc)
    expect = [oracle.edge_first_collision(b.robot, osc, b.q1, b.dir, b.K, b.step) for b in batches]
    for w in (4, 32):
        for spl in (1, 2, 4):
            engine.set_edge_tuning(w, spl)
            got = engine.edge_batches(eid, batches)
            for g, e in zip(got, expect):
                assert np.array_equal(g, e), (w, spl)
    ids, q1, u, K = engine.stack_edges(batches)
    best = engine.autotune_edges(eid, ids, q1, u, K)
    print("edge autotune:", best)
    assert best["waves_per_cu"] in (4, 8, 16, 32) and best["slots_per_lane"] in (1, 2, 4)
    got = engine.edge_batches(eid, batches)
    for g, e in zip(got, expect):
        assert np.array_equal(g, e)
