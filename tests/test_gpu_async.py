"""Stream-ordered queries (rvs_distance_async, rvs_is_valid_async): results
on a caller's stream equal the blocking calls' bit for bit; calls on two
streams that share the context's scratch run in call order; a blocking call
made while asynchronous calls are in flight waits for them; arguments the
asynchronous forms do not take are rejected."""
import numpy as np
import pytest
import torch

from rpmplv2_amd import rvs, workloads as W

pytestmark = pytest.mark.gpu


def _outputs(n, L, dev):
    return (torch.full((n,), float("nan"), device=dev), torch.full((n, L), float("nan"), device=dev),
            torch.full((n,), 255, dtype=torch.uint8, device=dev))


@pytest.mark.parametrize("precision", [rvs.PREC_FP32_REFINE, rvs.PREC_REF32, rvs.PREC_FP64])
def test_distance_async_matches_blocking(engine, precision):
    dev = torch.device("cuda:0")
    w = W.workload2(1 << 17)
    rid, eid = engine.load_robot(w.robot), engine.create_env(w.scene)
    n, L = len(w.q), w.robot.num_links
    qa = torch.from_numpy(w.q).to(dev)
    qb = torch.from_numpy(np.ascontiguousarray(w.q[::-1])).to(dev)
    ref_a, ref_b = _outputs(n, L, dev), _outputs(n, L, dev)
    engine.distance_device(rid, eid, qa, *ref_a, precision=precision)
    engine.distance_device(rid, eid, qb, *ref_b, precision=precision)
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    out_a, out_b = _outputs(n, L, dev), _outputs(n, L, dev)
    done_a, done_b = torch.cuda.Event(), torch.cuda.Event()
    # two calls on two streams back to back: they share the scratch arena
    engine.distance_async(rid, eid, qa, *out_a, precision=precision, stream=s1, done=done_a)
    engine.distance_async(rid, eid, qb, *out_b, precision=precision, stream=s2, done=done_b)
    done_a.synchronize()
    done_b.synchronize()
    for got, ref in ((out_a, ref_a), (out_b, ref_b)):
        for g, r in zip(got, ref):
            assert torch.equal(g, r)
    # a blocking call right behind asynchronous ones (no wait in between)
    out_c = _outputs(n, L, dev)
    engine.distance_async(rid, eid, qa, *out_c, precision=precision, stream=s1)
    d, prof, ver, _ = engine.distance(rid, eid, w.q[:1000], precision=precision)
    s1.synchronize()
    for g, r in zip(out_c, ref_a):
        assert torch.equal(g, r)
    assert np.array_equal(d, ref_a[0][:1000].cpu().numpy())


def test_is_valid_async_matches_blocking(engine):
    dev = torch.device("cuda:0")
    w = W.workload2(1 << 16)
    rid, eid = engine.load_robot(w.robot), engine.create_env(w.scene)
    expect = engine.is_valid(rid, eid, w.q)
    q = torch.from_numpy(w.q).to(dev)
    valid = torch.full((len(w.q),), 255, dtype=torch.uint8, device=dev)
    s = torch.cuda.Stream()
    done = torch.cuda.Event()
    engine.is_valid_async(rid, eid, q, valid, stream=s, done=done)
    done.synchronize()
    assert np.array_equal(valid.cpu().numpy(), expect)


def test_async_rejects_host_arrays_and_timing(engine):
    w = W.workload2(64)
    rid, eid = engine.load_robot(w.robot), engine.create_env(w.scene)
    q = np.ascontiguousarray(w.q)
    d = np.zeros(len(q), np.float32)
    import ctypes as C
    st = rvs.lib().rvs_distance_async(engine._ctx, rid, eid, q.ctypes.data, len(q), d.ctypes.data, None, None, 0,
                                      None, None)
    assert st == rvs.RVS_EINVAL and b"device arrays" in rvs.lib().rvs_last_error(engine._ctx)
    dev = torch.device("cuda:0")
    qd, dd = torch.from_numpy(q).to(dev), torch.empty(len(q), device=dev)
    st = rvs.lib().rvs_distance_async(engine._ctx, rid, eid, C.c_void_p(qd.data_ptr()), len(q),
                                      C.c_void_p(dd.data_ptr()), None, None, rvs.F_DEVICE | rvs.F_TIMING, None, None)
    assert st == rvs.RVS_EINVAL


def test_async_empty_batch(engine):
    """An empty batch enqueues nothing but still records the caller's event."""
    dev = torch.device("cuda:0")
    w = W.workload2(8)
    rid, eid = engine.load_robot(w.robot), engine.create_env(w.scene)
    q = torch.empty((0, 6), device=dev)
    d = torch.empty(0, device=dev)
    done = torch.cuda.Event()
    engine.distance_async(rid, eid, q, d, precision=rvs.PREC_FP32_REFINE, stream=torch.cuda.Stream(), done=done)
    done.synchronize()
    assert engine.distance(rid, eid, w.q, precision=rvs.PREC_FP32_REFINE)[2].shape == (8,)


def test_underestimation_and_self_collision_async(engine):
    """The stream-ordered underestimation and self-collision queries equal
    the blocking calls bit for bit."""
    dev = torch.device("cuda:0")
    w = W.workload2(1 << 14)
    rid, eid = engine.load_robot(w.robot), engine.create_env(w.scene)
    q = w.q
    _, _, _, npo = engine.distance(rid, eid, q[:1], precision=rvs.PREC_REF32, want_nearest=True)
    ref = npo[0]
    d0, p0, s0 = engine.underestimation(rid, eid, q, ref)
    m0, f0 = engine.self_collision(rid, q)
    qd = torch.from_numpy(q).to(dev)
    refd = torch.from_numpy(np.ascontiguousarray(ref)).to(dev)
    n, L = len(q), w.robot.num_links
    d = torch.full((n,), float("nan"), device=dev)
    prof = torch.full((n, L), float("nan"), device=dev)
    st = torch.full((n,), 255, dtype=torch.uint8, device=dev)
    m = torch.full((n,), float("nan"), device=dev)
    f = torch.full((n,), 255, dtype=torch.uint8, device=dev)
    s = torch.cuda.Stream()
    done = torch.cuda.Event()
    engine.underestimation_async(rid, eid, qd, refd, d, prof, st, stream=s)
    engine.self_collision_async(rid, qd, m, f, stream=s, done=done)
    done.synchronize()
    assert np.array_equal(d.cpu().numpy(), d0) and np.array_equal(st.cpu().numpy(), s0)
    assert np.array_equal(prof.cpu().numpy(), p0)
    assert np.array_equal(m.cpu().numpy(), m0) and np.array_equal(f.cpu().numpy(), f0)
