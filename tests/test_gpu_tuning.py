"""Autotuner of the FP32 bulk pass: every launch shape (one lane per
configuration, lane refill, persistent chunks) and walk order gives the same
per-configuration results (they only change which lane evaluates a
configuration); the per-lane walks, which evaluate pairs without the case
tree, give the same verdicts and distances within FP32 rounding; and
rvs_autotune keeps one of the candidate shapes."""
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
    # bit-identical within a first-hit search: in obstacle order (variants
    # 0-2), among all candidates in walk order (3), from the hit on (4); the
    # searches evaluate different pairs, so their screening flags differ
    refs = {}
    form_of = lambda v: "ordered" if v < 3 else v
    for coh in (0, 5, 3):
        for variant in (0, 1, 2, 3, 4):
            for block in ((256, 512, 1024) if variant >= 3 else (64, 128, 256)):
                for per_cu in ((0,) if variant == 0 else (0, 1)):
                    engine.set_tuning(rid, eid, variant, block, per_cu, coh)
                    d, prof, ver, _ = engine.distance(rid, eid, w.q, precision=rvs.PREC_FP32_REFINE)
                    form = form_of(variant)
                    shape = (variant, block, per_cu, coh)
                    if form not in refs:
                        refs[form] = (d, prof, ver)
                    else:
                        ref = refs[form]
                        assert np.array_equal(ver, ref[2]), shape
                        assert np.array_equal(d, ref[0]) and np.array_equal(prof, ref[1]), shape
    # across forms: the same verdicts, distances and profiles within FP32 rounding
    d0, p0, v0 = refs["ordered"]
    for form in (3, 4):
        d1, p1, v1 = refs[form]
        assert np.array_equal(v0, v1), form
        assert np.abs(d0 - d1).max() <= 1e-6 and np.abs(p0 - p1).max() <= 1e-6, form
    best = engine.autotune(rid, eid, w.q)
    print(wl, "autotune:", best)
    assert best["variant"] in (0, 1, 2, 3, 4) and best["block"] in (64, 128, 256, 512, 1024) and best["ms"] > 0
    assert engine.get_tuning(rid, eid)["variant"] == best["variant"]
    d, prof, ver, _ = engine.distance(rid, eid, w.q, precision=rvs.PREC_FP32_REFINE)
    ref = refs[form_of(best["variant"])]
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


@pytest.mark.parametrize("case", ["xarm6_ragged", "xarm6_one_pose", "chain2", "chain1"])
def test_walk_order_covers_every_configuration(engine, case):
    """The walk-order sort hands every configuration to exactly one lane at
    every key width: outputs start as NaN (verdict 255) and must equal the
    unordered pass's.  Cases: a ragged batch (70,001 configurations, not a
    whole number of tiles), 65,536 copies of one pose (one bucket cut into
    four chunks), and 2- and 1-joint chains (up to 9 and 18 bits per
    joint)."""
    import torch
    rng = np.random.default_rng(7)
    if case.startswith("xarm6"):
        w = W.workload2(70001 if case == "xarm6_ragged" else 65536)
        rb, scene, q = w.robot, w.scene, w.q
        if case == "xarm6_one_pose":
            q = np.repeat(q[:1], len(q), axis=0)
        widths = (0, 1, 2, 3, 4, 5, 6, 18)
    else:
        nd = 2 if case == "chain2" else 1
        rb = R.chain([0.3] * nd, radius=0.02)
        scene = W._boxes(rng, 6, 0.05, 0.1, 0.8)
        q = rng.uniform(-4.0, 4.0, (50000, nd)).astype(np.float32)
        widths = (0, 3, 9, 18)
    rid = engine.load_robot(rb)
    eid = engine.create_env(scene)
    dev = torch.device("cuda:0")
    qd = torch.from_numpy(np.ascontiguousarray(q)).to(dev)
    n, L = qd.shape[0], rb.num_links
    ref = None
    for cb in widths:
        for variant in (0, 1, 2):
            engine.set_tuning(rid, eid, variant, 128, 0, cb)
            d = torch.full((n,), float("nan"), device=dev)
            prof = torch.full((n, L), float("nan"), device=dev)
            ver = torch.full((n,), 255, dtype=torch.uint8, device=dev)
            engine.distance_device(rid, eid, qd, d, prof, ver, precision=rvs.PREC_FP32_REFINE)
            got = (d.cpu().numpy(), prof.cpu().numpy(), ver.cpu().numpy())
            assert not np.isnan(got[0]).any() and not np.isnan(got[1]).any() and (got[2] != 255).all(), (cb, variant)
            if ref is None:
                ref = got
            else:
                assert np.array_equal(got[2], ref[2]), (cb, variant)
                assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), (cb, variant)
