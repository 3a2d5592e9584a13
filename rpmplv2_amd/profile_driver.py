"""Small driver for rocprofv3 runs of the distance kernels.

    python -m rpmplv2_amd.profile_driver [--workload 2|5|bur|traj|edgeself|edges] [--n N] [--calls K] [--calibrate]
                                         [--tuning VARIANT,BLOCK,BLOCKS_PER_CU,COHERENCE_BITS]

Builds the seeded workload, optionally runs a 256 MiB elementwise copy first
(the calibration point for FETCH_SIZE / WRITE_SIZE), then calls the engine K
times on device-resident inputs: rvs_distance per precision (workloads 2 and
5), generalized spines from one root with underestimation and real-distance
layers (bur: 65,536 targets), trajectory safety (traj: 8,192 straight
60-point trajectories), xArm6 edge self-collision (edgeself: 16,384
edges) or the first collision along the planner-start edges of the four
chains (edges: 4 x 4,096 edges in one persistent launch, starts judged by
the engine's own reference-arithmetic edge search, shape 8 waves per CU x
4 slots, the autotuner's choice).  --tuning records a launch shape for the FP32 bulk pass (as
rvs_set_tuning does), so that a profile can use the shape an autotuned run
chose.  Used as the profiled
program in `rocprofv3 --kernel-trace --stats -- python3 -m ...` and
`rocprofv3 --pmc ... -- python3 -m ...` runs.
"""
from __future__ import annotations

import argparse


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="2")
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--calls", type=int, default=3)
    ap.add_argument("--calibrate", action="store_true")
    ap.add_argument("--precisions", default="ref32,fp32r,fp64")
    ap.add_argument("--tuning", default="")
    args = ap.parse_args()

    import torch
    from . import rvs, workloads as W

    dev = torch.device("cuda:0")
    if args.calibrate:
        src = torch.empty(64 << 20, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        del src, dst

    if args.workload in ("bur", "traj", "edgeself"):
        _bubble(args, torch, rvs, W, dev)
        return
    if args.workload == "edges":
        _edges(args, torch, rvs, W, dev)
        return
    w = W.workload2(args.n) if args.workload == "2" else W.workload5(args.n)
    eng = rvs.Engine()
    rid = eng.load_robot(w.robot)
    eid = eng.create_env(w.scene, rvs.ENV_OBB if args.workload == "5" else rvs.ENV_AABB)
    if args.tuning:
        v, b, bpc, cb = (int(x) for x in args.tuning.split(","))
        eng.set_tuning(rid, eid, v, b, bpc, cb)
    n, L = len(w.q), w.robot.num_links
    q = torch.from_numpy(w.q).to(dev)
    d = torch.empty(n, device=dev)
    prof = torch.empty((n, L), device=dev)
    ver = torch.empty(n, dtype=torch.uint8, device=dev)
    names = {"ref32": rvs.PREC_REF32, "fp32r": rvs.PREC_FP32_REFINE, "fp64": rvs.PREC_FP64}
    for name in args.precisions.split(","):
        for _ in range(args.calls):
            eng.distance_device(rid, eid, q, d, prof, ver, precision=names[name])
    torch.cuda.synchronize()
    print("profile driver done", flush=True)
    eng.close()


def _bubble(args, torch, rvs, W, dev):
    import numpy as np
    eng = rvs.Engine()
    if args.workload == "bur":
        for real in (False, True):
            w = W.workload6(1 << 16, use_real_dist=real)
            rid, eid = eng.load_robot(w.robot), eng.create_env(w.scene)
            qr, qe = torch.from_numpy(w.root).to(dev), torch.from_numpy(w.q_e).to(dev)
            qn = torch.empty_like(qe)
            st = torch.empty(len(qe), dtype=torch.int32, device=dev)
            dd = torch.empty(len(qe), device=dev)
            for _ in range(args.calls):
                eng.extend_gen_spines_device(rid, eid, qr, qe, qn, st, dd, num_layers=w.num_layers,
                                             num_iter=w.num_iter, use_real_dist=real, d_crit=w.d_crit)
    elif args.workload == "traj":
        w = W.workload2(64)
        keep = w.scene.is_ground == 0
        scene = W.Scene(w.scene.dim[keep], w.scene.pos[keep], w.scene.quat[keep], w.scene.is_ground[keep],
                        w.scene.min_dist_tol[keep])
        rid, eid = eng.load_robot(w.robot), eng.create_env(scene)
        root = np.zeros(6, np.float32)
        _, p0, _, np0 = eng.distance(rid, eid, root[None], want_nearest=True)
        rng = np.random.default_rng(82)
        u = rng.normal(size=(8192, 6))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        s = np.linspace(0, 1, 61)[1:]
        pts = (rng.uniform(0.1, 2.5, 8192)[:, None, None] * s[None, :, None] * u[:, None, :]).astype(np.float32)
        mv = np.full(scene.num_obj, 0.1, np.float32)
        for _ in range(args.calls):
            eng.trajectory_is_safe(rid, eid, pts, root, p0[0], np0[0], mv, self_collision=True)
    else:
        from . import robots as R
        rb = R.xarm6(radius=0.05)
        rng = np.random.default_rng(61)
        lim = np.asarray(rb.limits, np.float64)
        q1 = (lim[:, 0] + rng.random((16384, 6)) * (lim[:, 1] - lim[:, 0])).astype(np.float32)
        u = rng.normal(size=q1.shape)
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        q2 = (q1 + rng.uniform(0.05, 3.0, len(q1))[:, None] * u).astype(np.float32)
        rid = eng.load_robot(rb)
        for _ in range(args.calls):
            eng.edge_self_collision(rid, q1, q2, 0.1)
    torch.cuda.synchronize()
    print("profile driver done", flush=True)
    eng.close()


def _edges(args, torch, rvs, W, dev):
    import numpy as np
    eng = rvs.Engine()
    loaded = {}

    def judge(rb, scene, q):
        if id(scene) not in loaded:
            loaded[id(scene)] = eng.create_env(scene)
        if id(rb) not in loaded:
            loaded[id(rb)] = eng.load_robot(rb)
        first = eng.edge_first_collision(loaded[id(scene)], np.full(len(q), loaded[id(rb)], np.int32), q,
                                         np.zeros_like(q), np.ones(len(q), np.int32))
        return first < 0

    w = W.workload3_planner(judge=judge)
    eid = eng.create_env(w.scene)
    ids, q1, u, K = eng.stack_edges(w.batches)
    q1d, ud = torch.from_numpy(q1).to(dev), torch.from_numpy(u).to(dev)
    first = torch.empty(len(ids), dtype=torch.int32, device=dev)
    eng.set_edge_tuning(8, 4)
    for _ in range(args.calls):
        eng.edge_first_collision_device(eid, ids, q1d, ud, K, first)
    torch.cuda.synchronize()
    print("profile driver done", flush=True)
    eng.close()


if __name__ == "__main__":
    main()
