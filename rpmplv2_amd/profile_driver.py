"""Small driver for rocprofv3 runs of the distance kernels.

    python -m rpmplv2_amd.profile_driver [--workload 2|5] [--n N] [--calls K] [--calibrate]

Builds the seeded workload, optionally runs a 256 MiB elementwise copy first
(the calibration point for FETCH_SIZE / WRITE_SIZE), then calls rvs_distance
K times per precision on device-resident inputs.  Used as the profiled
program in `rocprofv3 --kernel-trace --stats -- python3 -m ...` and
`rocprofv3 --pmc ... -- python3 -m ...` runs.
"""
from __future__ import annotations

import argparse


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--calls", type=int, default=3)
    ap.add_argument("--calibrate", action="store_true")
    ap.add_argument("--precisions", default="ref32,fp32r,fp64")
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

    w = W.workload2(args.n) if args.workload == 2 else W.workload5(args.n)
    eng = rvs.Engine()
    rid = eng.load_robot(w.robot)
    eid = eng.create_env(w.scene, rvs.ENV_OBB if args.workload == 5 else rvs.ENV_AABB)
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


if __name__ == "__main__":
    main()
