#!/usr/bin/env python3
"""Benchmark: distance queries per second, xArm6 with 16 boxes + ground.

One step = one rvs_distance call over the whole batch of 1,048,576 joint
configurations (xArm6, capsule radius 0.05 m, ground box + 16 random boxes):
per-link distance profile, minimum distance and collision verdict for every
configuration, inputs and outputs resident in HBM.  The timed precision is
the engine's reference-exact arithmetic (verdicts bit-identical to RPMPLv2);
the FP64 path is measured beside it.

Prints one JSON line (rank 0).  Only one GPU is supported.
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
FLOP_FK_JOINT = 100
FLOP_CAPSULE_AABB = 160
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def workload_flops(rb, scene) -> int:
    pairs = 0
    for i in range(rb.num_links):
        for j in range(scene.num_obj):
            if scene.is_ground[j] and i < rb.ground_included:
                continue
            pairs += 1
    return rb.num_dofs * FLOP_FK_JOINT + pairs * FLOP_CAPSULE_AABB


def cpu_baseline(w, n_sample: int):
    """The oracle's reference arithmetic (ref32) timed on this host."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    import numpy as np
    sc = O.scene_of(w.scene)
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = max(1, min(cores, 16))
    q = w.q[:n_sample]
    O.distance(w.robot, sc, q[:4096], 32, nthreads=cores)  # warm-up
    runs = []
    for _ in range(3):
        t = time.perf_counter()
        O.distance(w.robot, sc, q, 32, nthreads=cores)
        runs.append(time.perf_counter() - t)
    t_all = sorted(runs)[1]
    q1 = q[:32768]
    t = time.perf_counter()
    O.distance(w.robot, sc, q1, 32, nthreads=1)
    t_one = time.perf_counter() - t
    return {"value": len(q) / t_all, "unit": "configs/s", "cores": cores, "kind": "port",
            "sample": f"{len(q)} configurations of the timed workload, median of 3 passes on {cores} threads",
            "value_1thread": len(q1) / t_one}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only; multi-GPU runs are not supported by this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from rpmplv2_amd import rvs, workloads as W

    w = W.workload2(args.n)
    n = len(w.q)
    eng = rvs.Engine(device=0, precision=rvs.PREC_REF32)
    rid = eng.load_robot(w.robot)
    eid = eng.create_env(w.scene)
    L = w.robot.num_links
    dev = torch.device("cuda:0")
    q = torch.from_numpy(w.q).to(dev)
    d = torch.empty(n, dtype=torch.float32, device=dev)
    prof = torch.empty((n, L), dtype=torch.float32, device=dev)
    ver = torch.empty(n, dtype=torch.uint8, device=dev)
    stream = torch.cuda.ExternalStream(eng.stream)
    torch.cuda.synchronize()

    def run(prec):
        eng.distance_device(rid, eid, q, d, prof, ver, precision=prec)

    for _ in range(args.warmup):
        run(rvs.PREC_REF32)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        evs[k][0].record(stream)
        run(rvs.PREC_REF32)
        evs[k][1].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    kern_ms = sum(a.elapsed_time(b) for a, b in evs) / args.steps
    value = n / (ms_step * 1e-3)

    # FP64 path beside it (same inputs)
    for _ in range(2):
        run(rvs.PREC_FP64)
    torch.cuda.synchronize()
    t = time.perf_counter()
    reps = max(3, args.steps // 4)
    for _ in range(reps):
        run(rvs.PREC_FP64)
    torch.cuda.synchronize()
    fp64_rate = n * reps / (time.perf_counter() - t)

    flop_cfg = workload_flops(w.robot, w.scene)
    achieved_tflops = flop_cfg * n / (kern_ms * 1e-3) / 1e12
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
        "data": "synthetic (seeded xArm6 configurations uniform in joint limits, 16 random boxes + ground)",
        "config": {"workload": "xarm6_16aabb_ground", "configs_per_step": n, "links": L,
                   "objects": w.scene.num_obj, "precision": "ref32 (FP64 kinematics, reference-order FP32 geometry)",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": achieved_tflops, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tflops / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "k_distance<float>", "kernel_ms": kern_ms, "flop_per_config": flop_cfg,
                     "hbm_GBps_algorithmic": hbm_bytes / (kern_ms * 1e-3) / 1e9},
        "fp64_configs_per_s": fp64_rate,
    }
    if not args.no_cpu:
        cb = cpu_baseline(w, min(n, 1 << 20))
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu_all"] = value / cb["value"]
    print(json.dumps(out))
    eng.close()


if __name__ == "__main__":
    main()
