"""rvs_distance_async_static, the entry point that allocates nothing: its
results equal rvs_distance_async's bit for bit (the same kernels; scratch
from the library's fixed device array instead of the context's arena) on two
calls in a row, the second with a smaller batch, in every precision; batches
beyond its limit are refused."""
import numpy as np
import pytest

from rpmplv2_amd import rvs, workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [rvs.PREC_FP32_REFINE, rvs.PREC_REF32, rvs.PREC_FP64])
def test_static_scratch_matches_async(prec):
    import torch
    eng = rvs.Engine(device=0)
    try:
        w = W.workload2(1 << 20)
        rid = eng.load_robot(w.robot)
        eid = eng.create_env(w.scene)
        dev = torch.device("cuda:0")
        L = w.robot.num_links
        for n in (1 << 20, 70001):          # the largest batch, then a smaller one
            q = torch.from_numpy(np.ascontiguousarray(w.q[-n:])).to(dev)
            outs = []
            for call in (eng.distance_async, eng.distance_async_static):
                d = torch.full((n,), float("nan"), device=dev)
                prof = torch.full((n, L), float("nan"), device=dev)
                ver = torch.full((n,), 255, dtype=torch.uint8, device=dev)
                call(rid, eid, q, d, prof, ver, precision=prec)
                torch.cuda.synchronize()
                outs.append((d.cpu().numpy(), prof.cpu().numpy(), ver.cpu().numpy()))
            (d0, p0, v0), (d1, p1, v1) = outs
            print(f"prec {prec} n {n}: free {v1.mean():.3f}")
            assert (v1 != 255).all() and not np.isnan(d1).any()
            assert np.array_equal(v0, v1) and np.array_equal(d0, d1) and np.array_equal(p0, p1)
        q = torch.zeros(((1 << 20) + 1, 6), device=dev)
        d = torch.empty((1 << 20) + 1, device=dev)
        with pytest.raises(rvs.RvsError):
            eng.distance_async_static(rid, eid, q, d, precision=prec)
    finally:
        eng.close()
