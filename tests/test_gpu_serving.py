"""The request server (rvs_server_*): continuous batching on one GPU.

Every request's result must equal what the direct entry points compute for
it alone against the obstacle positions of the tick the server used
(rvs_request_info.tick): verdicts and stored flags identical, distances equal
(the same kernels run on the same inputs; the stated tolerance is 1e-5 m).
"""
import threading

import numpy as np
import pytest

from rpmplv2_amd import rvs, workloads as W

pytestmark = pytest.mark.gpu


def _direct(engine, rid, scene, kind, q, cache):
    eid = cache.get(id(scene))
    if eid is None:
        eid = engine.create_env(scene)
        cache[id(scene)] = eid
    if kind == rvs.REQ_VERDICT:
        d, prof, ver, _ = engine.distance(rid, eid, q, precision=rvs.PREC_FP32_REFINE)
        return d, ver, prof
    _, _, _, npo = engine.distance(rid, eid, q[:1], precision=rvs.PREC_REF32, want_nearest=True)
    d, prof, st = engine.underestimation(rid, eid, q, npo[0])
    return d, st, prof


def _check_requests(engine, w, robot_ids, results):
    """results: list of (robot, kind, q, d, flag, prof, info)."""
    scenes = {}
    cache = {}
    mism = 0
    for robot, kind, q, d, flag, prof, info in results:
        tick = info["tick"]
        sc = w.scene if tick < 0 else scenes.setdefault(tick, W.scene_at_tick(w.scene, w.vel, tick, w.tick_seconds))
        dd, ff, pp = _direct(engine, robot_ids[robot], sc, kind, q, cache)
        mism += int((ff != flag).sum())
        assert np.array_equal(ff, flag), (robot, kind, len(q), tick)
        assert np.abs(dd - d).max() <= 1e-5
        if prof is not None:
            ok = flag == 1
            assert np.abs(pp[ok] - prof[ok]).max(initial=0) <= 1e-5
    return mism


def test_server_mixed_requests_match_direct_calls(engine):
    w = W.workload4(duration_s=0.05, pool=1 << 16)
    rids = [engine.load_robot(rb) for rb in w.robots]
    eid = engine.create_env(w.scene)
    srv = rvs.Server(engine, memory_budget=256 << 20, tick_seconds=w.tick_seconds)
    srv.set_motion(eid, w.vel)
    rng = np.random.default_rng(7)
    K = 240
    sizes = np.minimum(np.floor(np.exp(rng.uniform(0, np.log(8192), K))).astype(int), 8192)
    reqs = []
    for k in range(K):
        r = int(rng.integers(0, 2))
        kind = int(rng.integers(0, 2))
        off = int(rng.integers(0, len(w.pools[r]) - sizes[k]))
        reqs.append((r, kind, w.pools[r][off:off + sizes[k]], bool(k % 3 == 0)))
    out = [None] * K

    def planner(p):
        mine = [k for k in range(K) if k % 4 == p]
        tickets = [(k, srv.submit(rids[reqs[k][0]], eid, reqs[k][1], reqs[k][2], want_profile=reqs[k][3]))
                   for k in mine]
        for k, t in tickets:
            out[k] = srv.wait(t)

    th = [threading.Thread(target=planner, args=(p,)) for p in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    st = srv.stats()
    srv.close()
    print("server stats:", st)
    assert st["completed"] == K and st["configs_completed"] == int(sizes.sum())
    batched = sum(1 for o in out if o[3]["batch_requests"] > 1)
    print(f"{batched} of {K} requests shared a launch; ticks used {sorted({o[3]['tick'] for o in out})[:5]}...")
    assert batched > 0
    results = [(reqs[k][0], reqs[k][1], reqs[k][2], out[k][0], out[k][1], out[k][2], out[k][3]) for k in range(K)]
    assert _check_requests(engine, w, rids, results) == 0


def _ulps(a, b):
    ia = np.ascontiguousarray(a, np.float32).view(np.int32).astype(np.int64)
    ib = np.ascontiguousarray(b, np.float32).view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    return np.abs(ia - ib)


def _check_against_oracle(oracle, w, results):
    """Every served result against the oracle on the obstacle positions of
    its tick.  Verdict requests (FP32 + refinement): verdicts identical to
    the reference arithmetic, distances and profiles within 1e-5 m of it and
    of the FP64 twin where the two agree on the verdict.  Underestimation
    requests (reference arithmetic, nearest points of the request's first
    configuration, RealVectorSpace.cpp:318-392 then :402-446): stored flags
    identical, distances and stored profiles within 2 float ulps (the
    device's and the host's double sine and cosine may differ in the last
    bit)."""
    scenes = {}
    for robot, kind, q, d, flag, prof, info in results:
        tick = info["tick"]
        if tick not in scenes:
            sc = w.scene if tick < 0 else W.scene_at_tick(w.scene, w.vel, tick, w.tick_seconds)
            scenes[tick] = oracle.scene_of(sc)
        osc = scenes[tick]
        rb = w.robots[robot]
        what = (robot, kind, len(q), tick)
        if kind == rvs.REQ_VERDICT:
            od, oprof, over, _ = oracle.distance(rb, osc, q, 32)
            d64, _, v64, _ = oracle.distance(rb, osc, q, 64)
            assert np.array_equal(flag, over), what
            assert np.abs(d - od).max() <= 1e-5, what
            agree = over == v64
            assert np.abs(d - d64)[agree].max(initial=0) <= 1e-5, what
            if prof is not None:
                assert np.abs(prof - oprof).max() <= 1e-5, what
        else:
            _, _, _, onp = oracle.distance(rb, osc, q[:1], 32, want_np=True)
            ud, uprof, ust = oracle.underestimation(rb, osc, q, onp[0])
            assert np.array_equal(flag, ust), what
            assert np.array_equal(np.isnan(d), np.isnan(ud)), what
            m = ~np.isnan(d)
            assert _ulps(d[m], ud[m]).max(initial=0) <= 2, what
            if prof is not None:
                ok = ust == 1
                assert _ulps(prof[ok], uprof[ok]).max(initial=0) <= 2, what


def test_server_spec_sizes_match_oracle(oracle, engine):
    """The serving workload's request sizes, floor(exp U[0, ln 65536]) up to
    65,536 configurations, both robots (xArm6, 10-DOF chain) and both kinds
    (verdict, underestimation), from four planner threads with the obstacles
    moving every 10 ms.  Requests above 16,384 configurations are copied
    directly instead of through the server's pinned staging, so both copy
    paths of both kinds run.  Each result is checked against the oracle on
    the obstacle positions of the tick the server used, and against a direct
    call of the engine made for the request alone (the same result in the mix
    and alone)."""
    w = W.workload4(duration_s=0.05, pool=1 << 17)
    rids = [engine.load_robot(rb) for rb in w.robots]
    eid = engine.create_env(w.scene)
    srv = rvs.Server(engine, memory_budget=512 << 20, tick_seconds=w.tick_seconds)
    srv.set_motion(eid, w.vel)
    rng = np.random.default_rng(19)
    K = 128
    sizes = np.clip(np.floor(np.exp(rng.uniform(0, np.log(65536), K))).astype(int), 1, 65536)
    robot = rng.integers(0, 2, K)
    kind = rng.integers(0, 2, K)
    # the largest sizes on every (robot, kind): the direct-copy paths
    for i, (r, k) in enumerate([(0, 0), (0, 1), (1, 0), (1, 1)]):
        sizes[2 * i], sizes[2 * i + 1] = 65536, 16385 + 7919 * i
        robot[2 * i] = robot[2 * i + 1] = r
        kind[2 * i] = kind[2 * i + 1] = k
    reqs = []
    for k in range(K):
        r = int(robot[k])
        off = int(rng.integers(0, len(w.pools[r]) - sizes[k] + 1))
        reqs.append((r, int(kind[k]), w.pools[r][off:off + sizes[k]], bool(k % 2 == 0)))
    out = [None] * K

    def planner(p):
        mine = [k for k in range(K) if k % 4 == p]
        tickets = [(k, srv.submit(rids[reqs[k][0]], eid, reqs[k][1], reqs[k][2], want_profile=reqs[k][3]))
                   for k in mine]
        for k, t in tickets:
            out[k] = srv.wait(t)

    th = [threading.Thread(target=planner, args=(p,)) for p in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    st = srv.stats()
    srv.close()
    print("server stats:", st)
    assert st["completed"] == K and st["configs_completed"] == int(sizes.sum())
    ticks = sorted({o[3]["tick"] for o in out})
    print(f"{int(sizes.sum())} configurations in {K} requests; {int((sizes > 16384).sum())} above the staging "
          f"limit; {sum(1 for o in out if o[3]['batch_requests'] > 1)} shared a launch; ticks {ticks[:3]}..{ticks[-1]}")
    results = [(reqs[k][0], reqs[k][1], reqs[k][2], out[k][0], out[k][1], out[k][2], out[k][3]) for k in range(K)]
    _check_against_oracle(oracle, w, results)
    assert _check_requests(engine, w, rids, results) == 0


def test_server_memory_budget_queues(engine):
    """A budget that holds about one request at a time: requests wait for
    memory, none is lost, device use never exceeds the budget."""
    w = W.workload4(duration_s=0.01, pool=1 << 16)
    rid = engine.load_robot(w.robots[0])
    eid = engine.create_env(w.scene)
    srv = rvs.Server(engine, memory_budget=1 << 20, tick_seconds=w.tick_seconds)
    q = w.pools[0][:24000]
    # about 0.66 MB of device state each: one at a time fits the budget
    tickets = [srv.submit(rid, eid, rvs.REQ_VERDICT, q[k * 1000:k * 1000 + 20000]) for k in range(4)]
    res = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    assert st["peak_bytes"] <= 1 << 20
    assert st["batches"] == 4 and st["max_inflight_batches"] == 1
    results = [(0, 0, q[k * 1000:k * 1000 + 20000], r[0], r[1], r[2], r[3]) for k, r in enumerate(res)]
    _check_requests(engine, w, [rid], results)


def test_server_edge_cases(engine):
    w = W.workload4(duration_s=0.01, pool=1 << 12)
    rid = engine.load_robot(w.robots[0])
    eid = engine.create_env(w.scene)
    srv = rvs.Server(engine, memory_budget=1 << 20)
    t = srv.submit(rid, eid, rvs.REQ_VERDICT, w.pools[0][:0])
    d, flag, prof, info = srv.wait(t)
    assert len(d) == 0
    with pytest.raises(rvs.RvsError):
        srv.submit(rid, eid + 100, rvs.REQ_VERDICT, w.pools[0][:4])
    with pytest.raises(rvs.RvsError):      # larger than the 1 MiB budget
        srv.submit(rid, eid, rvs.REQ_VERDICT, np.zeros((40000, 6), np.float32))
    t = srv.submit(rid, eid, rvs.REQ_UNDERESTIMATION, w.pools[0][:1])
    assert srv.wait(t)[3]["tick"] == -1      # no motion set: the static scene
    srv.close()
