"""Autotuner of the FP32 bulk pass: every launch shape gives the same
per-configuration results (the shape only changes which lane evaluates a
configuration), and rvs_autotune keeps one of the candidate shapes."""
import numpy as np
import pytest

from rpmplv2_amd import rvs, workloads as W

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
