#!/usr/bin/env python3
"""Benchmark: distance queries per second, xArm6 with 16 boxes + ground.

Headline step = one rvs_distance call over the whole batch of 1,048,576
joint configurations (xArm6, capsule radius 0.05 m, ground box + 16 random
boxes, configurations uniform in the joint limits): per-link distance
profile, minimum distance and collision verdict for every configuration,
inputs and outputs resident in HBM.  The timed path is the FP32 bulk pass
with exact near-contact refinement, whose verdicts are identical to the
reference arithmetic; the exact path and the FP64 path are timed beside it.
After timing, the whole batch is checked against the CPU oracle (verdict
mismatches against the reference arithmetic, distance error against the
FP64 twin), and the oracle's reference arithmetic is timed on the host cores
as the CPU baseline.

Further sections of the same JSON line (skip with --quick):
  workload5  22-DOF chain, 512 rotated boxes, 16,777,216 configurations:
             oriented boxes and the reference's AABB-of-OBB reading, self-
             collision, error against the FP64 oracle on a 65,536-configuration
             stride-256 subsample
  edges      chains of 10/14/18/22 DOF, 4,096 discretised edges each, first
             colliding state per edge in one persistent launch
  serving    Poisson request stream at 20,000 requests/s for 5 s through the
             request server (p50 / p99 latency, throughput)

Prints one JSON line.  One GPU only.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic work per unit (fixed nominal counts used for the roofline):
# revolute FK joint 100 flop, capsule-AABB pair 160, capsule-OBB pair 196
# (the box-frame transform of the two end points), capsule-capsule pair 60,
# and the broad-phase test of a pair (bounding-box gap against the running
# link minimum) 15.  The kernels skip the pairs the broad phase rules out, so
# a launch processes: FK joints, one broad-phase test per (link, obstacle)
# pair reached before the early exit, and the pairs that pass it.  Those
# counts come from the oracle's restatement of the same loop
# (orc_distance_cull32) on the same configurations.
FLOP_FK_JOINT = 100
FLOP_CAPSULE_AABB = 160
FLOP_CAPSULE_OBB = 196
FLOP_CAPSULE_CAPSULE = 60
FLOP_BROAD_PHASE = 15
PEAK_FP32_TFLOPS = 157.3
# the FP32 vector rate without packed FP32 instructions, which this build
# turns off (one FP32 FMA per lane per cycle instead of two)
PEAK_FP32_UNPACKED_TFLOPS = 78.6
PEAK_HBM_GBPS = 8000.0
# FP32 operations per configuration that the bulk kernels execute, from
# rocprofv3 counters on the same workloads: SQ_INSTS_VALU_FLOPS_FP32 (2 FMA +
# ADD + MUL + TRANS, in wave instructions) x lanes active per VALU
# instruction (SQ_THREAD_CYCLES_VALU / SQ_ACTIVE_INST_VALU), divided by the
# configurations of the profiled call.  roofline.achieved = this x the
# configurations of a launch / the launch time measured here.
COUNTER_FLOP_PER_CONFIG = {
    ("xarm6_16aabb_ground", "k_distance_fast"): (2276.6, "profiles/k_distance_fast_xarm6_1M.md"),
    ("chain22_512rotated_boxes", "k_distance_fast_links"):
        (15056.3, "profiles/k_distance_fast_links_chain22_512obb.md"),
}
# VALU wave instructions per configuration of the same kernels (SQ_INSTS_VALU
# / configurations, same profiles): the issue rate they reach against the
# chip's 614.4 G wave instructions/s (1,024 SIMDs x 2.4 GHz / 4 cycles per
# wave64 instruction), the limit these compare- and select-heavy kernels meet
# long before the FP32 peak.
COUNTER_VALU_PER_CONFIG = {
    ("xarm6_16aabb_ground", "k_distance_fast"): 74.77,
    ("chain22_512rotated_boxes", "k_distance_fast_links"): 775.31,
}
PEAK_VALU_ISSUE_G = 1024 * 2.4 / 4
# HBM bytes per configuration of the link-major bulk pass on the 22-DOF chain
# with 512 rotated boxes (all 22 launches of a call): FETCH_SIZE 1,730,503.6
# KiB x 2.0 + WRITE_SIZE 1,900,112.3 KiB x 1.0 per call on 2,097,152
# configurations (profiles/k_distance_fast_links_chain22_512obb.md).
LINKS_TRAFFIC_PER_CONFIG = (2.0 * 1733610.1 + 1900112.3) * 1024 / 2097152
# The first-collision search on the planner-start edges (four chains, 4,096
# edges each, shape 8 waves per CU x 4 slots), per launch, from rocprofv3
# counters (profiles/k_edge_first_chains_planner_starts.md): VALU wave
# instructions, FP32 and FP64 operations executed (SQ_INSTS_VALU_FLOPS_FP32 /
# _FP64 x 26.11 lanes active per VALU instruction) and HBM bytes (FETCH_SIZE
# 38,058 KiB x 2.0 + WRITE_SIZE 2,540.1 KiB).
EDGE_COUNTERS = {"valu_instr": 2.7824e8, "fp32_flop": 1.0554e8 * 26.11, "fp64_flop": 3.0155e7 * 26.11,
                 "traffic_bytes": (2.0 * 38058.0 + 2540.1) * 1024,
                 "source": "profiles/k_edge_first_chains_planner_starts.md"}
PEAK_FP64_TFLOPS = 78.6
# HBM bytes per launch of the FP32 bulk kernel on the headline workload from
# two `rocprofv3 --pmc` runs: FETCH_SIZE 161,910.5 KiB x 2.0 + WRITE_SIZE
# 124,679.7 KiB x 1.0, the factors being what the same counters read for a
# 256 MiB copy (profiles/k_distance_fast_xarm6_1M.md).
TRAFFIC_BYTES = (2.0 * 161910.5 + 124679.7) * 1024   # rocprofv3 FETCH_SIZE x 2 + WRITE_SIZE per launch
BULK_KERNELS = {0: "k_distance_fast", 1: "k_distance_fast_pool", 2: "k_distance_fast_chunks", 3: "k_distance_fast_lanes",
                4: "k_distance_fast_links"}


def valu_roofline(workload: str, kernel: str, n: int, kernel_ms: float, evaluated_flop_cfg: float,
                  dense_flop_cfg: float) -> dict:
    """roofline of an FP32 bulk kernel: achieved = FP32 operations executed
    per configuration (counters, when the kernel has a profile on this
    workload; else the evaluated-work model) x n / kernel time."""
    entry = COUNTER_FLOP_PER_CONFIG.get((workload, kernel))
    if entry:
        flop_cfg, src = entry
        accounting = ("FP32 operations executed per configuration from rocprofv3 counters (" + src +
                      "): SQ_INSTS_VALU_FLOPS_FP32 x lanes active per VALU instruction")
    else:
        flop_cfg = evaluated_flop_cfg
        accounting = "evaluated-work model (no counter profile of this kernel on this workload)"
    achieved = flop_cfg * n / (kernel_ms * 1e-3) / 1e12
    valu_cfg = COUNTER_VALU_PER_CONFIG.get((workload, kernel))
    issue = None
    if valu_cfg:
        rate = valu_cfg * n / (kernel_ms * 1e-3) / 1e9
        issue = {"valu_instr_per_config": valu_cfg, "achieved_G_per_s": rate, "peak_G_per_s": PEAK_VALU_ISSUE_G,
                 "frac": rate / PEAK_VALU_ISSUE_G,
                 "note": "VALU wave instructions issued per second (counters x configurations / kernel time)"}
    return {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "valu_issue": issue,
            "frac": achieved / PEAK_FP32_TFLOPS, "kernel": kernel, "kernel_ms": kernel_ms,
            "flop_per_config": flop_cfg, "flop_accounting": accounting,
            "peak_unpacked": PEAK_FP32_UNPACKED_TFLOPS, "frac_unpacked": achieved / PEAK_FP32_UNPACKED_TFLOPS,
            "peak_kind": "FP32 vector (VALU) with packed FP32; peak_unpacked without it, as this build runs; "
                         "the path has no MFMA work",
            "evaluated_flop_per_config": evaluated_flop_cfg,
            "evaluated_tflops": evaluated_flop_cfg * n / (kernel_ms * 1e-3) / 1e12,
            "evaluated_note": "model: 100 flop per FK joint + 15 per broad-phase test + 160 (AABB) or 196 (OBB) per "
                              "evaluated pair, the counts from the oracle's restatement of the walk",
            "dense_flop_per_config": dense_flop_cfg,
            "dense_equivalent_tflops": dense_flop_cfg * n / (kernel_ms * 1e-3) / 1e12,
            "dense_note": "nominal work of every FK joint and every (link, obstacle) pair per second; the broad "
                          "phase skips most pairs, so it is no achieved rate"}


def pair_count(rb, scene) -> int:
    pairs = 0
    for i in range(rb.num_links):
        for j in range(scene.num_obj):
            if scene.is_ground[j] and i < rb.ground_included:
                continue
            pairs += 1
    return pairs


# Host threads for the CPU baseline: the CPU share of one GPU on the
# benchmark machines is 16 cores, whatever the process's affinity reports.
HOST_THREAD_CAP = 16


def host_info() -> dict:
    """The host the CPU baseline runs on: CPU model, logical CPUs, the
    process's affinity and the threads used."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"cpu_model": model, "nproc": os.cpu_count() or 1, "affinity": affinity,
            "threads": max(1, min(affinity, HOST_THREAD_CAP)),
            "thread_cap": f"{HOST_THREAD_CAP} (the CPU share of one GPU on this machine class)"}


def host_cores() -> int:
    return host_info()["threads"]


def oracle_module():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    return O


def cpu_baseline(w, cores: int):
    """The oracle's reference arithmetic (ref32) timed on this host: a fresh
    state per query, FK + skeleton + pair loop with the early exit."""
    O = oracle_module()
    sc = O.scene_of(w.scene)
    q = w.q
    O.distance(w.robot, sc, q[:4096], 32, nthreads=cores)  # warm-up
    runs, res = [], None
    for _ in range(3):
        t = time.perf_counter()
        res = O.distance(w.robot, sc, q, 32, nthreads=cores)
        runs.append(time.perf_counter() - t)
    t_all = sorted(runs)[1]
    q1 = q[:65536]
    t = time.perf_counter()
    O.distance(w.robot, sc, q1, 32, nthreads=1)
    t_one = time.perf_counter() - t
    cb = {"value": len(q) / t_all, "unit": "configs/s", "cores": cores, "kind": "port",
          "sample": f"all {len(q)} configurations of the timed batch, median of 3 passes on {cores} threads "
                    f"({t_all:.2f} s each); 1-thread rate on the first 65536",
          "value_1thread": len(q1) / t_one, "host": host_info()}
    return cb, res


def timed_distance(eng, rid, eid, q, d, prof, ver, prec, steps, warmup):
    """Wall time per call over `steps` calls as a planner would make them
    (the FP32 path replays its captured HIP graph), then the kernel stage
    times from HIP events over max(3, steps // 4) more calls."""
    import numpy as np
    import torch
    for _ in range(warmup):
        eng.distance_device(rid, eid, q, d, prof, ver, precision=prec)
    torch.cuda.synchronize()
    refined = []
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.distance_device(rid, eid, q, d, prof, ver, precision=prec)
        refined.append(eng.last_refined())
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    bulk, refine, order = [], [], []
    for _ in range(max(3, steps // 4) if steps > 1 else 1):
        eng.distance_device(rid, eid, q, d, prof, ver, precision=prec, timing=True)
        st = eng.last_timing()
        bulk.append(st[0])
        refine.append(st[1] if len(st) > 1 else 0.0)
        order.append(st[2] if len(st) > 2 else 0.0)
    torch.cuda.synchronize()
    timed_distance.order_ms = float(np.mean(order))
    return ms, float(np.mean(bulk)), float(np.mean(refine)), int(np.mean(refined))


def workload5_section(eng, cores: int, n: int) -> dict:
    import numpy as np
    import torch
    from rpmplv2_amd import rvs, workloads as W
    O = oracle_module()
    w = W.workload5(n)
    rb = w.robot
    L = rb.num_links
    rid = eng.load_robot(rb)
    eo = eng.create_env(w.scene, rvs.ENV_OBB)
    ea = eng.create_env(w.scene, rvs.ENV_AABB)
    dev = torch.device("cuda:0")
    q = torch.from_numpy(w.q).to(dev)
    d = torch.empty(n, device=dev)
    prof = torch.empty((n, L), device=dev)
    ver = torch.empty(n, dtype=torch.uint8, device=dev)
    tune = eng.autotune(rid, eo, q)
    ms_o, bulk_o, ref_o, nref_o = timed_distance(eng, rid, eo, q, d, prof, ver, rvs.PREC_FP32_REFINE, 2, 1)
    sub = np.arange(0, n, 256)
    d_o, ver_o, free_o = d.cpu().numpy()[sub], ver.cpu().numpy()[sub], float(ver.float().mean().item())
    ms_a, bulk_a, ref_a, nref_a = timed_distance(eng, rid, ea, q, d, prof, ver, rvs.PREC_FP32_REFINE, 3, 1)
    d_a, ver_a, free_a = d.cpu().numpy()[sub], ver.cpu().numpy()[sub], float(ver.float().mean().item())
    ms_r32, k_r32, _, _ = timed_distance(eng, rid, eo, q, d, prof, ver, rvs.PREC_REF32, 1, 0)
    d64 = torch.empty(n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    t = time.perf_counter()
    eng.distance_f64_device(rid, eo, q, d64, None, ver)
    torch.cuda.synchronize()
    ms_f64 = (time.perf_counter() - t) * 1e3
    smin = torch.empty(n, device=dev)
    sfree = torch.empty(n, dtype=torch.uint8, device=dev)
    eng.self_collision_device(rid, q, smin, sfree)
    torch.cuda.synchronize()
    t = time.perf_counter()
    eng.self_collision_device(rid, q, smin, sfree)
    torch.cuda.synchronize()
    ms_self = (time.perf_counter() - t) * 1e3
    smin_s, sfree_all = smin.cpu().numpy()[sub], float(sfree.float().mean().item())
    # the stated configuration's verdict: free of the boxes (oriented, FP32 +
    # refinement) and of itself, two calls back to back on the same batch
    torch.cuda.synchronize()
    t = time.perf_counter()
    eng.distance_device(rid, eo, q, d, prof, ver, precision=rvs.PREC_FP32_REFINE)
    eng.self_collision_device(rid, q, smin, sfree)
    torch.cuda.synchronize()
    ms_both = (time.perf_counter() - t) * 1e3
    both = (ver.cpu().numpy() == 1) & (sfree.cpu().numpy() == 1)
    both_s = both[sub]
    del q, d, prof, ver, d64, smin, sfree
    torch.cuda.empty_cache()
    # oracle on the stride-256 subsample
    qs = w.q[sub]
    so, sa = O.scene_of(w.scene, oriented=True), O.scene_of(w.scene)
    od32, _, ov32, _ = O.distance(rb, so, qs, 32, nthreads=cores)
    od64, _, ov64, _ = O.distance(rb, so, qs, 64, nthreads=cores)
    ad32, _, av32, _ = O.distance(rb, sa, qs, 32, nthreads=cores)
    osm = O.self_min(rb, qs, nthreads=cores)
    _, _, _, (ev_p, tests_p) = O.distance_cull(rb, so, qs, 32, nthreads=cores)
    m = len(qs)
    flop_obb_dense = rb.num_dofs * FLOP_FK_JOINT + L * w.scene.num_obj * FLOP_CAPSULE_OBB
    flop_obb = rb.num_dofs * FLOP_FK_JOINT + (tests_p * FLOP_BROAD_PHASE + ev_p * FLOP_CAPSULE_OBB) / m
    flop_aabb = rb.num_dofs * FLOP_FK_JOINT + L * w.scene.num_obj * FLOP_CAPSULE_AABB
    bytes_cfg = rb.num_dofs * 4 + 4 + L * 4 + 1
    return {
        "workload": "chain22_512rotated_boxes", "configs": n, "self_pairs": len(rb.self_pairs),
        "obb": {"configs_per_s": n / (ms_o * 1e-3), "ms": ms_o, "bulk_ms": bulk_o, "refine_ms": ref_o,
                "refined": nref_o, "collision_free_fraction": free_o, "tuning": tune,
                "verdict_mismatch_vs_ref32_subsample": int((ver_o != ov32).sum()),
                "max_abs_dist_err_vs_ref64_subsample": float(np.abs(d_o - od64).max()),
                "roofline": dict(valu_roofline("chain22_512rotated_boxes", BULK_KERNELS[tune["variant"]], n, bulk_o,
                                               flop_obb, flop_obb_dense),
                                 pairs_evaluated_per_config=ev_p / m, box_tests_per_config=tests_p / m,
                                 counts_from="oracle restatement of the bulk pass's walk on the "
                                             "65,536-configuration subsample",
                                 algorithmic_bytes_per_config=bytes_cfg,
                                 hbm_GBps_algorithmic=bytes_cfg * n / (bulk_o * 1e-3) / 1e9,
                                 **({"traffic_bytes_per_config": LINKS_TRAFFIC_PER_CONFIG,
                                     "traffic_GBps": LINKS_TRAFFIC_PER_CONFIG * n / (bulk_o * 1e-3) / 1e9,
                                     "traffic_note": "FETCH_SIZE x 2 + WRITE_SIZE over the 22 link launches, per "
                                                     "configuration (profiles/k_distance_fast_links_chain22_512obb.md)"}
                                    if tune["variant"] == 4 else {}))},
        "aabb_of_obb_reference_semantics": {
            "configs_per_s": n / (ms_a * 1e-3), "ms": ms_a, "bulk_ms": bulk_a, "refine_ms": ref_a,
            "refined": nref_a, "collision_free_fraction": free_a,
            "verdict_mismatch_vs_ref32_subsample": int((ver_a != av32).sum()),
            "max_abs_dist_err_vs_ref32_subsample": float(np.abs(d_a - ad32).max()),
            "nominal_flop_per_config": flop_aabb},
        "obb_ref32_exact": {"configs_per_s": n / (ms_r32 * 1e-3), "kernel_ms": k_r32},
        "obb_fp64": {"configs_per_s": n / (ms_f64 * 1e-3), "ms": ms_f64},
        "self_collision": {"configs_per_s": n / (ms_self * 1e-3), "ms": ms_self, "free_fraction": sfree_all,
                           "max_abs_min_err_vs_ref32_subsample": float(np.abs(smin_s - osm).max())},
        "verdict_with_self_collision": {
            "configs_per_s": n / (ms_both * 1e-3), "ms": ms_both,
            "what": "oriented-box distance (FP32 + refinement) and self-collision (210 pairs) of every "
                    "configuration, two calls back to back; verdict = free of both",
            "free_fraction": float(both.mean()),
            "verdict_mismatch_vs_ref32_subsample": int((both_s != ((ov32 == 1) & (osm >= 0))).sum())},
        "subsample": "65,536 configurations (every 256th) checked against the oracle",
    }


def workload1_section(eng, cores: int) -> dict:
    """Planar 2-DOF arm, 4 boxes, 65,536 configurations: the FP64 drop-in
    (rvs_distance_f64 on device arrays), the reference arithmetic and the
    FP32 path, with the oracle's FP64 loop on the host cores beside them."""
    import numpy as np
    import torch
    from rpmplv2_amd import rvs, workloads as W
    O = oracle_module()
    w = W.workload1(65536)
    rid, eid = eng.load_robot(w.robot), eng.create_env(w.scene)
    n, L = len(w.q), w.robot.num_links
    dev = torch.device("cuda:0")
    q = torch.from_numpy(w.q).to(dev)
    d64 = torch.empty(n, dtype=torch.float64, device=dev)
    p64 = torch.empty((n, L), dtype=torch.float64, device=dev)
    v = torch.empty(n, dtype=torch.uint8, device=dev)
    for _ in range(3):
        eng.distance_f64_device(rid, eid, q, d64, p64, v)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(20):
        eng.distance_f64_device(rid, eid, q, d64, p64, v)
    torch.cuda.synchronize()
    ms64 = (time.perf_counter() - t0) * 1e3 / 20
    got_d, got_v = d64.cpu().numpy(), v.cpu().numpy()
    d = torch.empty(n, device=dev)
    prof = torch.empty((n, L), device=dev)
    ms32, _, _, _ = timed_distance(eng, rid, eid, q, d, prof, v, rvs.PREC_REF32, 20, 3)
    msr, _, _, _ = timed_distance(eng, rid, eid, q, d, prof, v, rvs.PREC_FP32_REFINE, 20, 3)
    sc = O.scene_of(w.scene)
    O.distance(w.robot, sc, w.q[:4096], 64, nthreads=cores)
    t = time.perf_counter()
    od, _, ov, _ = O.distance(w.robot, sc, w.q, 64, nthreads=cores)
    cpu_s = time.perf_counter() - t
    return {"workload": "planar2dof_4boxes", "configs": n,
            "fp64": {"configs_per_s": n / (ms64 * 1e-3), "ms": ms64,
                     "verdict_mismatch_vs_oracle_fp64": int((got_v != ov).sum()),
                     "max_abs_dist_err_vs_oracle_fp64": float(np.abs(got_d - od).max())},
            "ref32": {"configs_per_s": n / (ms32 * 1e-3), "ms": ms32},
            "fp32_refine": {"configs_per_s": n / (msr * 1e-3), "ms": msr},
            "cpu_oracle_fp64": {"configs_per_s": n / cpu_s, "cores": cores}}


def _edge_run(eng, O, w, cores: int, profiled: bool = False) -> dict:
    """One persistent first-collision launch over every edge of the four
    chains (after rvs_autotune_edges), checked against the oracle.  profiled:
    the workload of the counter profile (EDGE_COUNTERS), which then gives
    the achieved rates over this run's kernel time."""
    import numpy as np
    import torch
    eid = eng.create_env(w.scene)
    ids, q1, u, K = eng.stack_edges(w.batches)
    dev = torch.device("cuda:0")
    q1d, ud = torch.from_numpy(q1).to(dev), torch.from_numpy(u).to(dev)
    first = torch.empty(len(ids), dtype=torch.int32, device=dev)
    tune = eng.autotune_edges(eid, ids, q1d, ud, K)
    eng.edge_first_collision_device(eid, ids, q1d, ud, K, first)
    torch.cuda.synchronize()
    # each call returns after its stream is synchronized; ms is the median
    # wall time of 11 calls (host planning and launches included), so that a
    # one-off stall of the host does not stand for the rate
    ks, walls = [], []
    for _ in range(11):
        t0 = time.perf_counter()
        eng.edge_first_collision_device(eid, ids, q1d, ud, K, first, timing=True)
        walls.append((time.perf_counter() - t0) * 1e3)
        ks.append(eng.last_timing()[0])
    torch.cuda.synchronize()
    ms = float(np.median(walls))
    kms = float(np.median(ks))
    got = first.cpu().numpy()
    sc = O.scene_of(w.scene)
    mism, off, states_eval, flop_eval = 0, 0, 0, 0.0
    for b in w.batches:
        m = len(b.q1)
        of = O.edge_first_collision(b.robot, sc, b.q1, b.dir, b.K, b.step, nthreads=cores)
        g = got[off:off + m]
        mism += int((g != of).sum())
        ev = np.where(g >= 0, g + 1, b.K).astype(np.int64)       # states the edge's search reaches
        n = b.robot.num_dofs
        flop_state = (n * FLOP_FK_JOINT + n * w.scene.num_obj * FLOP_CAPSULE_AABB
                      + len(b.robot.self_pairs) * FLOP_CAPSULE_CAPSULE)
        states_eval += int(ev.sum())
        flop_eval += float(ev.sum()) * flop_state
        off += m
    nominal = int(K.sum())
    dense = flop_eval / (kms * 1e-3) / 1e12
    roof = {"bound": "valu", "kernel": "k_edge_first_q", "achieved": None, "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "dense_equivalent_tflops": dense, "dense_frac": dense / PEAK_FP32_TFLOPS,
            "dense_note": "nominal work of every state a search reaches (edge start to its first collision or "
                          "its end): 100 per FK joint + 160 per capsule-box pair + 60 per self pair; the axis "
                          "tests and early exits evaluate far less"}
    if profiled:
        C = EDGE_COUNTERS
        sec = kms * 1e-3
        fp32 = C["fp32_flop"] / sec / 1e12
        roof.update({"achieved": fp32, "frac": fp32 / PEAK_FP32_TFLOPS,
                     "flop_accounting": "FP32 operations executed per launch from rocprofv3 counters (" +
                                        C["source"] + ") / this run's kernel time",
                     "fp64": {"achieved": C["fp64_flop"] / sec / 1e12, "peak": PEAK_FP64_TFLOPS,
                              "note": "the forward kinematics in double, as the reference's KDL chain"},
                     "valu_issue": {"valu_instr_per_launch": C["valu_instr"],
                                    "achieved_G_per_s": C["valu_instr"] / sec / 1e9, "peak_G_per_s": PEAK_VALU_ISSUE_G,
                                    "frac": C["valu_instr"] / sec / 1e9 / PEAK_VALU_ISSUE_G},
                     "traffic": C["traffic_bytes"]})
    return {"edges": len(ids), "states_nominal": nominal, "states_evaluated": states_eval,
            "ms": ms, "kernel_ms": kms, "edges_per_s": len(ids) / (ms * 1e-3),
            "states_evaluated_per_s": states_eval / (ms * 1e-3),
            "states_nominal_per_s": nominal / (ms * 1e-3), "tuning": tune,
            "first_collision_mismatch_vs_oracle": mism,
            "edges_colliding_at_state0": int((got == 0).sum()), "edges_free": int((got == -1).sum()),
            "roofline": roof}


def edges_section(eng, cores: int) -> dict:
    """Chains of 10/14/18/22 DOF, 4,096 edges each, one persistent launch:
    as specified (every start collides: one of the 64 boxes overlaps the
    chains' fixed base link) and with starts drawn collision-free, as a
    tree planner's nodes are (workload3_planner)."""
    from rpmplv2_amd import workloads as W
    O = oracle_module()
    out = {"workload": "chains_10_14_18_22dof_64boxes_4096edges_each"}
    out["as_specified"] = _edge_run(eng, O, W.workload3(), cores)
    out["planner_starts"] = _edge_run(eng, O, W.workload3_planner(oracle=O), cores, profiled=True)
    return out


def is_valid_section(eng, cores: int) -> dict:
    """RealVectorSpace::isValid(q) (the collision tree) on the headline
    batch, device arrays, reference arithmetic and FP64, checked against the
    oracle, which is also the CPU baseline."""
    import numpy as np
    import torch
    from rpmplv2_amd import workloads as W
    O = oracle_module()
    w = W.workload2()
    n = len(w.q)
    rid = eng.load_robot(w.robot)
    eid = eng.create_env(w.scene)
    qd = torch.from_numpy(w.q).to(torch.device("cuda:0"))
    out = torch.empty(n, dtype=torch.uint8, device=qd.device)
    sc = O.scene_of(w.scene)
    res = {"workload": "xarm6_16aabb_ground", "configs": n}
    from rpmplv2_amd import rvs
    for name, prec, bits in (("ref32", rvs.PREC_REF32, 32), ("fp64", rvs.PREC_FP64, 64)):
        eng.is_valid(rid, eid, qd, prec, out)
        torch.cuda.synchronize()
        ks = []
        for _ in range(5):
            eng.is_valid(rid, eid, qd, prec, out, timing=True)
            ks.append(eng.last_timing())
        t0 = time.perf_counter()
        for _ in range(5):
            eng.is_valid(rid, eid, qd, prec, out)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / 5
        got = out.cpu().numpy()
        O.is_valid(w.robot, sc, w.q[:4096], bits, nthreads=cores)   # warm-up
        t0 = time.perf_counter()
        ov = O.is_valid(w.robot, sc, w.q, bits, nthreads=cores)
        cpu_s = time.perf_counter() - t0
        res[name] = {"configs_per_s": n / (ms * 1e-3), "ms": ms, "kernel_ms": float(np.min([k[0] for k in ks])),
                     "walk_order_ms": float(np.min([k[1] for k in ks])), "kernel": "k_valid",
                     "valid_fraction": float(got.mean()), "mismatch_vs_oracle": int((got != ov).sum()),
                     "cpu_baseline": {"value": n / cpu_s, "unit": "configs/s", "cores": cores, "kind": "port",
                                      "sample": "the same 1,048,576 configurations"}}
    return res


def underestimation_section(eng, cores: int) -> dict:
    """RealVectorSpace::computeDistanceUnderestimation on the headline batch
    against one nearest-point set (the oracle's, of the first collision-free
    configuration), device arrays, checked against the oracle, which is also
    the CPU baseline."""
    import numpy as np
    import torch
    from rpmplv2_amd import workloads as W
    O = oracle_module()
    w = W.workload2()
    n, L = len(w.q), w.robot.num_links
    sc = O.scene_of(w.scene)
    _, _, over, onp = O.distance(w.robot, sc, w.q[:64], 32, want_np=True, nthreads=cores)
    ref = np.ascontiguousarray(onp[int(np.argmax(over == 1))])
    rid = eng.load_robot(w.robot)
    eid = eng.create_env(w.scene)
    dev = torch.device("cuda:0")
    qd, refd = torch.from_numpy(w.q).to(dev), torch.from_numpy(ref).to(dev)
    d = torch.empty(n, device=dev)
    prof = torch.empty((n, L), device=dev)
    st = torch.empty(n, dtype=torch.uint8, device=dev)
    eng.underestimation_device(rid, eid, qd, refd, d, prof, st)
    torch.cuda.synchronize()
    ks = []
    for _ in range(5):
        eng.underestimation_device(rid, eid, qd, refd, d, prof, st, timing=True)
        ks.append(eng.last_timing()[0])
    t0 = time.perf_counter()
    for _ in range(5):
        eng.underestimation_device(rid, eid, qd, refd, d, prof, st)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / 5
    gd, gst = d.cpu().numpy(), st.cpu().numpy()
    O.underestimation(w.robot, sc, w.q[:4096], ref, nthreads=cores)   # warm-up
    t0 = time.perf_counter()
    ud, _, ust = O.underestimation(w.robot, sc, w.q, ref, nthreads=cores)
    cpu_s = time.perf_counter() - t0
    same = gst == ust
    ia, ib = gd.view(np.int32).astype(np.int64), ud.view(np.int32).astype(np.int64)
    ulp = int(np.abs(np.where(ia < 0, -(ia & 0x7FFFFFFF), ia) - np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)).max())
    return {"workload": "xarm6_16aabb_ground", "configs": n, "ms": ms, "kernel_ms": float(np.min(ks)),
            "configs_per_s": n / (ms * 1e-3), "kernel": "k_under", "stored_fraction": float(gst.mean()),
            "stored_mismatch_vs_oracle": int((~same).sum()), "max_ulp_d_vs_oracle": ulp,
            "cpu_baseline": {"value": n / cpu_s, "unit": "configs/s", "cores": cores, "kind": "port",
                             "sample": "the same 1,048,576 configurations"}}


def edge_validity_section(eng, cores: int) -> dict:
    """RealVectorSpace::isValid(q1, q2) for 65,536 short xArm6 edges in the
    headline scene (10 interpolated states each, k = 10..1), device arrays,
    checked edge by edge against the oracle, which is also the CPU baseline."""
    import numpy as np
    import torch
    from rpmplv2_amd import workloads as W
    O = oracle_module()
    w = W.workload2(1 << 17)
    n = 1 << 16
    q1 = np.ascontiguousarray(w.q[:n])
    q2 = (q1 + 0.05 * (w.q[n:2 * n] - q1)).astype(np.float32)
    rid = eng.load_robot(w.robot)
    eid = eng.create_env(w.scene)
    dev = torch.device("cuda:0")
    q1d, q2d = torch.from_numpy(q1).to(dev), torch.from_numpy(q2).to(dev)
    out = torch.empty(n, dtype=torch.uint8, device=dev)
    eng.edge_is_valid(rid, eid, q1d, q2d, out)
    torch.cuda.synchronize()
    ks = []
    for _ in range(5):
        eng.edge_is_valid(rid, eid, q1d, q2d, out, timing=True)
        ks.append(eng.last_timing()[0])
    t0 = time.perf_counter()
    for _ in range(5):
        eng.edge_is_valid(rid, eid, q1d, q2d, out)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / 5
    got = out.cpu().numpy()
    sc = O.scene_of(w.scene)
    O.edge_valid(w.robot, sc, q1[:4096], q2[:4096], nthreads=cores)   # warm-up
    t0 = time.perf_counter()
    ov = O.edge_valid(w.robot, sc, q1, q2, nthreads=cores)
    cpu_s = time.perf_counter() - t0
    kms = float(np.min(ks))
    return {"workload": "xarm6_16aabb_ground_65536_short_edges_10_checks", "edges": n, "states": n * 10,
            "ms": ms, "kernel_ms": kms, "edges_per_s": n / (ms * 1e-3), "states_per_s": n * 10 / (ms * 1e-3),
            "valid_fraction": float(got.mean()), "mismatch_vs_oracle": int((got != ov).sum()),
            "kernel": "k_edge_valid_q",
            "cpu_baseline": {"value": n / cpu_s, "unit": "edges/s", "cores": cores, "kind": "port",
                             "sample": "the same 65,536 edges"}}


def spines_section(eng, cores: int) -> dict:
    """Generalized bur: extendGenSpine from one root towards 65,536 targets."""
    import numpy as np
    import torch
    from rpmplv2_amd import workloads as W
    O = oracle_module()
    out = {}
    for real in (False, True):
        w = W.workload6(1 << 16, use_real_dist=real)
        n = len(w.q_e)
        rid = eng.load_robot(w.robot)
        eid = eng.create_env(w.scene)
        dev = torch.device("cuda:0")
        qr, qe = torch.from_numpy(w.root).to(dev), torch.from_numpy(w.q_e).to(dev)
        qn = torch.empty_like(qe)
        st = torch.empty(n, dtype=torch.int32, device=dev)
        dd = torch.empty(n, dtype=torch.float32, device=dev)
        kw = dict(num_layers=w.num_layers, num_iter=w.num_iter, use_real_dist=w.use_real_dist, d_crit=w.d_crit)
        for _ in range(2):
            eng.extend_gen_spines_device(rid, eid, qr, qe, qn, st, dd, **kw)
        torch.cuda.synchronize()
        calls, t0 = 10, time.perf_counter()
        for _ in range(calls):
            eng.extend_gen_spines_device(rid, eid, qr, qe, qn, st, dd, **kw)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / calls
        g_qn, g_st, g_d = qn.cpu().numpy(), st.cpu().numpy(), dd.cpu().numpy()
        sc = O.scene_of(w.scene)
        m = 8192
        t1 = time.perf_counter()
        o_qn, o_st, o_d = O.extend_gen_spines(w.robot, sc, w.root, w.q_e[:m], w.num_layers, w.num_iter, True,
                                              w.use_real_dist, w.d_crit, nthreads=cores)
        cpu_s = time.perf_counter() - t1
        mism = int((~((g_qn[:m] == o_qn).all(1) & (g_st[:m] == o_st) & (g_d[:m] == o_d))).sum())
        out["real_distance" if real else "underestimation"] = {
            "workload": w.name, "spines": n, "num_layers": w.num_layers, "num_iter_spine": w.num_iter,
            "d_crit": w.d_crit, "ms": ms, "spines_per_s": n / (ms * 1e-3),
            "status_counts": {"advanced": int((g_st == 1).sum()), "trapped": int((g_st == 2).sum()),
                              "reached": int((g_st == 3).sum())},
            "mismatch_vs_oracle": mism, "checked": m,
            "cpu_baseline": {"value": m / cpu_s, "unit": "spines/s", "cores": cores, "kind": "port",
                             "sample": f"first {m} targets of the same bur"}}
        del qr, qe, qn, st, dd
    return out


def bubble_section(eng, cores: int) -> dict:
    """Trajectory safety (8,192 straight 60-point trajectories, obstacles at
    0.1 m/s) and xArm6 edge self-collision (16,384 edges), timed on the GPU
    and on the oracle."""
    import numpy as np
    from rpmplv2_amd import robots as R, workloads as W
    O = oracle_module()
    out = {}
    w = W.workload2(64)
    keep = w.scene.is_ground == 0
    scene = W.Scene(w.scene.dim[keep], w.scene.pos[keep], w.scene.quat[keep], w.scene.is_ground[keep],
                    w.scene.min_dist_tol[keep])
    rid, eid = eng.load_robot(w.robot), eng.create_env(scene)
    root = np.zeros(6, np.float32)
    _, p0, _, np0 = eng.distance(rid, eid, root[None], want_nearest=True)
    rng = np.random.default_rng(82)
    T, K = 8192, 60
    u = rng.normal(size=(T, 6))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    s = np.linspace(0, 1, K + 1)[1:]
    pts = (rng.uniform(0.1, 2.5, T)[:, None, None] * s[None, :, None] * u[:, None, :]).astype(np.float32)
    mv = np.full(scene.num_obj, 0.1, np.float32)
    eng.trajectory_is_safe(rid, eid, pts, root, p0[0], np0[0], mv, self_collision=True)
    t0 = time.perf_counter()
    for _ in range(5):
        safe = eng.trajectory_is_safe(rid, eid, pts, root, p0[0], np0[0], mv, self_collision=True)
    ms = (time.perf_counter() - t0) * 1e3 / 5
    sc = O.scene_of(scene)
    t1 = time.perf_counter()
    osafe = O.traj_safe(w.robot, sc, pts, root, p0[0], np0[0], mv, pairs=w.robot.self_pairs, nthreads=cores)
    cpu_s = time.perf_counter() - t1
    out["trajectories"] = {"workload": "xarm6_16aabb_8192traj_60pts", "trajectories": T, "points": K,
                           "ms": ms, "trajectories_per_s": T / (ms * 1e-3), "safe": int(safe.sum()),
                           "mismatch_vs_oracle": int((safe != osafe).sum()),
                           "cpu_baseline": {"value": T / cpu_s, "unit": "trajectories/s", "cores": cores,
                                            "kind": "port", "sample": "the same 8192 trajectories"},
                           "note": "host arrays: includes the 11.8 MB point upload"}
    rb = R.xarm6(radius=0.05)
    lim = np.asarray(rb.limits, np.float64)
    rng = np.random.default_rng(61)
    q1 = (lim[:, 0] + rng.random((16384, 6)) * (lim[:, 1] - lim[:, 0])).astype(np.float32)
    u = rng.normal(size=q1.shape)
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    q2 = (q1 + rng.uniform(0.05, 3.0, len(q1))[:, None] * u).astype(np.float32)
    rid = eng.load_robot(rb)
    eng.edge_self_collision(rid, q1, q2, 0.1)
    t0 = time.perf_counter()
    for _ in range(5):
        col, qo = eng.edge_self_collision(rid, q1, q2, 0.1)
    ms = (time.perf_counter() - t0) * 1e3 / 5
    t1 = time.perf_counter()
    ocol, oqo = O.edge_self_collision(rb, q1, q2, 0.1, nthreads=cores)
    cpu_s = time.perf_counter() - t1
    out["edge_self_collision"] = {"workload": "xarm6_16384edges_0.05-3rad_eps0.1", "edges": len(q1), "ms": ms,
                                  "edges_per_s": len(q1) / (ms * 1e-3), "colliding": int(col.sum()),
                                  "mismatch_vs_oracle": int(((col != ocol) | (qo != oqo).any(1)).sum()),
                                  "cpu_baseline": {"value": len(q1) / cpu_s, "unit": "edges/s", "cores": cores,
                                                   "kind": "port", "sample": "the same 16384 edges"},
                                  "note": "host arrays: includes the uploads and read-back"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--w5-n", type=int, default=1 << 24)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the oracle checks")
    ap.add_argument("--quick", action="store_true", help="headline workload only")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only; this build does not run across GPUs", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from rpmplv2_amd import rvs, workloads as W

    w = W.workload2(args.n)
    n = len(w.q)
    L = w.robot.num_links
    eng = rvs.Engine(device=0, precision=rvs.PREC_FP32_REFINE)
    rid = eng.load_robot(w.robot)
    eid = eng.create_env(w.scene)
    dev = torch.device("cuda:0")
    q = torch.from_numpy(w.q).to(dev)
    d = torch.empty(n, dtype=torch.float32, device=dev)
    prof = torch.empty((n, L), dtype=torch.float32, device=dev)
    ver = torch.empty(n, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    tune = eng.autotune(rid, eid, q)

    ms_step, bulk_ms, refine_ms, refined = timed_distance(eng, rid, eid, q, d, prof, ver, rvs.PREC_FP32_REFINE,
                                                          args.steps, args.warmup)
    order_ms = timed_distance.order_ms
    value = n / (ms_step * 1e-3)
    d_gpu, prof_gpu, ver_gpu = d.cpu().numpy(), prof.cpu().numpy(), ver.cpu().numpy()
    ms_ref32, k_ref32, _, _ = timed_distance(eng, rid, eid, q, d, prof, ver, rvs.PREC_REF32, max(3, args.steps // 2), 2)
    ms_fp64, k_fp64, _, _ = timed_distance(eng, rid, eid, q, d, prof, ver, rvs.PREC_FP64, max(3, args.steps // 2), 2)
    # end to end from host arrays (pageable memory): the configurations go up
    # over PCIe and d, profile and verdict come back inside each call
    for _ in range(2):
        eng.distance(rid, eid, w.q, precision=rvs.PREC_FP32_REFINE)
    t0 = time.perf_counter()
    for _ in range(5):
        eng.distance(rid, eid, w.q, precision=rvs.PREC_FP32_REFINE)
    ms_h2h = (time.perf_counter() - t0) * 1e3 / 5
    del q, d, prof, ver
    torch.cuda.empty_cache()

    # work per launch: the FP32 operations the kernel executes (counters of
    # the same kernel on this workload) x the configurations of the launch;
    # beside it the evaluated-work model (pairs reaching the broad phase and
    # pairs evaluated, counted by the oracle's restatement of the bulk pass's
    # walk) and the nominal dense figure
    O = oracle_module()
    _, _, _, (ev_pairs, box_tests) = O.distance_cull(w.robot, O.scene_of(w.scene), w.q, 32, nthreads=host_cores())
    dense_flop_cfg = w.robot.num_dofs * FLOP_FK_JOINT + pair_count(w.robot, w.scene) * FLOP_CAPSULE_AABB
    eval_flop_cfg = (n * w.robot.num_dofs * FLOP_FK_JOINT + box_tests * FLOP_BROAD_PHASE
                     + ev_pairs * FLOP_CAPSULE_AABB) / n
    hbm_bytes = n * (w.robot.num_dofs * 4 + 4 + L * 4 + 1)
    roof = valu_roofline("xarm6_16aabb_ground", BULK_KERNELS[tune["variant"]], n, bulk_ms, eval_flop_cfg,
                         dense_flop_cfg)
    roof.update({"traffic": TRAFFIC_BYTES, "pairs_evaluated_per_config": ev_pairs / n,
                 "box_tests_per_config": box_tests / n, "algorithmic_bytes": hbm_bytes,
                 "hbm_GBps_algorithmic": hbm_bytes / (bulk_ms * 1e-3) / 1e9,
                 "traffic_note": "FETCH_SIZE x 2 + WRITE_SIZE per launch of k_distance_fast on this workload "
                                 "(profiles/k_distance_fast_xarm6_1M.md)"})
    out = {
        "metric": "distance queries/sec (xArm6, 16 boxes) at verdict-exact",
        "value": value,
        "unit": "configs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded xArm6 configurations uniform in the joint limits; ground box + 16 random boxes)",
        "config": {"workload": "xarm6_16aabb_ground", "configs_per_step": n, "links": L,
                   "objects": w.scene.num_obj, "capsule_radius": 0.05,
                   "precision": "fp32 bulk + exact near-contact refinement (band 1e-4 m)"},
        "roofline": roof,
        "tuning": tune,
        "refine": {"configs": refined, "fraction": refined / n,
                   "kernel": "k_distance_refine_pairs" if L * w.scene.num_obj <= 256 else "k_distance_refine",
                   "kernel_ms": refine_ms},
        "walk_order": {"kernels": "k_coh_hist, k_coh_scan, k_coh_bucket, k_coh_lowsort "
                                  "(sort by the first joint angles)",
                       "ms": order_ms},
        "ref32_exact": {"configs_per_s": n / (ms_ref32 * 1e-3), "kernel_ms": k_ref32},
        "fp64": {"configs_per_s": n / (ms_fp64 * 1e-3), "kernel_ms": k_fp64},
        "host_to_host": {"configs_per_s": n / (ms_h2h * 1e-3), "ms": ms_h2h,
                         "note": "FP32 + refinement from pageable host arrays through rvs_distance (24 B per "
                                 "configuration up, 29 B down over PCIe inside each call); not value"},
    }
    cores = host_cores()
    if not args.no_cpu:
        cb, ref = cpu_baseline(w, cores)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu_all"] = value / cb["value"]
        od, oprof, over, _ = ref
        O = oracle_module()
        d64, prof64, ver64, _ = O.distance(w.robot, O.scene_of(w.scene), w.q, 64, nthreads=cores)
        free = (ver_gpu == 1) & (ver64 == 1)
        out["parity"] = {
            "verdict_mismatch_vs_ref32": int((ver_gpu != over).sum()),
            "max_abs_dist_err_vs_ref64": float(np.abs(d_gpu - d64).max()),
            "max_abs_profile_err_vs_ref64_free": float(np.abs(prof_gpu[free] - prof64[free]).max()),
            "collision_free_fraction": float(ver_gpu.mean()),
        }
    if not args.quick:
        print("bench: workload 1", file=sys.stderr, flush=True)
        out["workload1"] = workload1_section(eng, cores)
        print("bench: workload 5", file=sys.stderr, flush=True)
        out["workload5"] = workload5_section(eng, cores, args.w5_n)
        print("bench: edges", file=sys.stderr, flush=True)
        out["edges"] = edges_section(eng, cores)
        print("bench: edge validity", file=sys.stderr, flush=True)
        out["edge_validity"] = edge_validity_section(eng, cores)
        print("bench: isValid", file=sys.stderr, flush=True)
        out["is_valid"] = is_valid_section(eng, cores)
        print("bench: underestimation", file=sys.stderr, flush=True)
        out["underestimation"] = underestimation_section(eng, cores)
        print("bench: spines", file=sys.stderr, flush=True)
        out["spines"] = spines_section(eng, cores)
        print("bench: trajectories, edge self-collision", file=sys.stderr, flush=True)
        out["bubble"] = bubble_section(eng, cores)
        eng.close()
        print("bench: serving", file=sys.stderr, flush=True)
        from rpmplv2_amd import serving
        out["serving"] = serving.run(5.0)
    else:
        eng.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
