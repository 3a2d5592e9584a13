"""Captured HIP graphs of the FP32 + refinement pipeline (device-array
rvs_distance calls): every replay equals the kernel-by-kernel result, with
eager work and synchronisations between replays and in-place input updates."""
import numpy as np
import pytest
import torch

from rpmplv2_amd import rvs, workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("variant", [0, 1, 3, 4])
def test_graph_replays_match_eager(engine, variant):
    w = W.workload2(1 << 16)
    rid = engine.load_robot(w.robot)
    eid = engine.create_env(w.scene)
    engine.set_tuning(rid, eid, variant, 256, 0)
    dev = torch.device("cuda:0")
    n, L = len(w.q), w.robot.num_links
    q = torch.from_numpy(w.q).to(dev)
    bufs = [torch.empty(n, device=dev), torch.empty((n, L), device=dev), torch.empty(n, dtype=torch.uint8, device=dev)]
    ref = [torch.empty_like(b) for b in bufs]
    other = W.workload2(1 << 17).q[-n:]

    def eager():
        engine.distance_device(rid, eid, q, *ref, precision=rvs.PREC_FP32_REFINE, eager=True)
        return [r.clone() for r in ref], engine.last_refined()

    expect, nref = eager()
    for rep in range(6):
        engine.distance_device(rid, eid, q, *bufs, precision=rvs.PREC_FP32_REFINE)
        assert engine.last_refined() == nref, rep
        for a, b in zip(bufs, expect):
            assert torch.equal(a, b), rep
        # eager work between replays: torch kernels, a host-array query, a sync
        x = torch.randn(1 << 20, device=dev)
        (x * 2).sum().item()
        engine.distance(rid, eid, w.q[:1000], precision=rvs.PREC_FP32_REFINE)
        torch.cuda.synchronize()
        if rep == 2:
            # new inputs in place: the graph reads the current memory
            q[: n // 2].copy_(torch.from_numpy(other[: n // 2]).to(dev))
            expect, nref = eager()
    engine.set_tuning(rid, eid, 0, 256, 0)
