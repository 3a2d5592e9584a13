#!/usr/bin/env python3
"""Benchmark: distance queries per second, xArm6 with 16 boxes + ground.

One step = one rvs_distance call over the whole batch of 1,048,576 joint
configurations (xArm6, capsule radius 0.05 m, ground box + 16 random boxes,
configurations uniform in the joint limits): per-link distance profile,
minimum distance and collision verdict for every configuration, inputs and
outputs resident in HBM.  The timed path is the FP32 bulk pass with exact
near-contact refinement, whose verdicts are identical to the reference
arithmetic; the exact path and the FP64 path are timed beside it.

After timing, the same batch is checked against the CPU oracle: verdict
mismatches against the reference arithmetic (must be 0) and the largest
distance error against the FP64 twin (must be <= 1e-5 m).

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
# revolute FK joint 100 flop, capsule-AABB pair 160 flop.
FLOP_FK_JOINT = 100
FLOP_CAPSULE_AABB = 160
PEAK_FP32_TFLOPS = 157.3
# HBM bytes per launch of k_distance_fast on this workload from two
# `rocprofv3 --pmc` runs: FETCH_SIZE 12,504.7 KiB x 2.0 + WRITE_SIZE
# 58,144.8 KiB x 1.0, the factors being what the same counters read for a
# 256 MiB copy (profiles/k_distance_fast_xarm6_1M.md).
TRAFFIC_BYTES = (2.0 * 12504.7 + 1.0 * 58144.8) * 1024


def workload_flops(rb, scene) -> int:
    pairs = 0
    for i in range(rb.num_links):
        for j in range(scene.num_obj):
            if scene.is_ground[j] and i < rb.ground_included:
                continue
            pairs += 1
    return rb.num_dofs * FLOP_FK_JOINT + pairs * FLOP_CAPSULE_AABB


def host_cores() -> int:
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, min(cores or 1, 16))


def cpu_baseline(w, cores: int):
    """The oracle's reference arithmetic (ref32) timed on this host: a fresh
    state per query, FK + skeleton + pair loop with the early exit."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
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
          "value_1thread": len(q1) / t_one}
    return cb, res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the oracle check")
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

    def timed(prec, steps, warmup):
        for _ in range(warmup):
            eng.distance_device(rid, eid, q, d, prof, ver, precision=prec)
        torch.cuda.synchronize()
        bulk, refine, refined = [], [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            eng.distance_device(rid, eid, q, d, prof, ver, precision=prec, timing=True)
            st = eng.last_timing()
            bulk.append(st[0])
            refine.append(st[1] if len(st) > 1 else 0.0)
            refined.append(eng.last_refined())
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        return ms, float(np.mean(bulk)), float(np.mean(refine)), int(np.mean(refined))

    ms_step, bulk_ms, refine_ms, refined = timed(rvs.PREC_FP32_REFINE, args.steps, args.warmup)
    value = n / (ms_step * 1e-3)
    d_gpu, prof_gpu, ver_gpu = d.cpu().numpy(), prof.cpu().numpy(), ver.cpu().numpy()
    ms_ref32, k_ref32, _, _ = timed(rvs.PREC_REF32, max(3, args.steps // 2), 2)
    ms_fp64, k_fp64, _, _ = timed(rvs.PREC_FP64, max(3, args.steps // 2), 2)

    flop_cfg = workload_flops(w.robot, w.scene)
    achieved = flop_cfg * n / (bulk_ms * 1e-3) / 1e12
    hbm_bytes = n * (w.robot.num_dofs * 4 + 4 + L * 4 + 1)
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
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": TRAFFIC_BYTES,
                     "kernel": "k_distance_fast", "kernel_ms": bulk_ms, "flop_per_config": flop_cfg,
                     "algorithmic_bytes": hbm_bytes,
                     "hbm_GBps_algorithmic": hbm_bytes / (bulk_ms * 1e-3) / 1e9},
        "refine": {"configs": refined, "fraction": refined / n, "kernel": "k_distance_refine",
                   "kernel_ms": refine_ms},
        "ref32_exact": {"configs_per_s": n / (ms_ref32 * 1e-3), "kernel_ms": k_ref32},
        "fp64": {"configs_per_s": n / (ms_fp64 * 1e-3), "kernel_ms": k_fp64},
    }
    if not args.no_cpu:
        cores = host_cores()
        cb, ref = cpu_baseline(w, cores)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu_all"] = value / cb["value"]
        od, oprof, over, _ = ref
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O
        d64, prof64, ver64, _ = O.distance(w.robot, O.scene_of(w.scene), w.q, 64, nthreads=cores)
        free = (ver_gpu == 1) & (ver64 == 1)
        out["parity"] = {
            "verdict_mismatch_vs_ref32": int((ver_gpu != over).sum()),
            "max_abs_dist_err_vs_ref64": float(np.abs(d_gpu - d64).max()),
            "max_abs_profile_err_vs_ref64_free": float(np.abs(prof_gpu[free] - prof64[free]).max()),
            "collision_free_fraction": float(ver_gpu.mean()),
        }
    print(json.dumps(out))
    eng.close()


if __name__ == "__main__":
    main()
