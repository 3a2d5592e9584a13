"""Serving workload: write a request stream for the replay driver and run it.

    python -m rpmplv2_amd.serving [--duration 5] [--rate 20000]

`write_problem` serialises a workloads.ServingWorkload (robots, obstacles with
velocities, configuration pools, request stream, server options) for
rpmplv2_amd/_build/rvs_serve_replay, the C++ program that replays the stream
in real time against the request server and prints latency and throughput as
one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import subprocess
import tempfile

import numpy as np

from . import rvs, workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
REPLAY = os.path.join(HERE, "_build", "rvs_serve_replay")


def write_problem(w: W.ServingWorkload, path: str, memory_budget: int = 8 << 30, num_streams: int = 2,
                  max_batch_configs: int = 1 << 20, planners: int = 8, launches_per_stream: int = 1) -> None:
    with open(path, "wb") as f:
        f.write(b"RVSW")
        f.write(np.int32(1).tobytes())
        f.write(np.int32(len(w.robots)).tobytes())
        for rb, pool in zip(w.robots, w.pools):
            desc, keep = rvs.robot_desc(rb)
            f.write(np.array([rb.skeleton_kind, len(rb.joints), rb.num_links, rb.ground_included,
                              len(rb.self_pairs), rb.num_dofs], np.int32).tobytes())
            f.write(np.float32(rb.gripper_length).tobytes())
            f.write(C.string_at(C.addressof(keep[0]), C.sizeof(keep[0])))
            f.write(np.ascontiguousarray(rb.radius, np.float32).tobytes())
            f.write(np.ascontiguousarray(np.asarray(rb.self_pairs, np.int32).reshape(-1)).tobytes())
            pool = np.ascontiguousarray(pool, np.float32)
            f.write(np.int64(len(pool)).tobytes())
            f.write(pool.tobytes())
        boxes = rvs.boxes_of(w.scene)
        f.write(np.int32(w.scene.num_obj).tobytes())
        f.write(C.string_at(C.addressof(boxes), C.sizeof(boxes)))
        f.write(np.ascontiguousarray(w.vel, np.float32).tobytes())
        f.write(np.float32(w.tick_seconds).tobytes())
        f.write(np.int64(len(w.arrival)).tobytes())
        f.write(np.ascontiguousarray(w.arrival, np.float64).tobytes())
        f.write(np.ascontiguousarray(w.robot, np.int32).tobytes())
        f.write(np.ascontiguousarray(w.kind, np.int32).tobytes())
        f.write(np.ascontiguousarray(w.size, np.int64).tobytes())
        f.write(np.ascontiguousarray(w.offset, np.int64).tobytes())
        f.write(np.int64(memory_budget).tobytes())
        f.write(np.array([num_streams, max_batch_configs, planners, launches_per_stream], np.int32).tobytes())


def run(duration_s: float = 5.0, rate: float = 20000.0, memory_budget: int = 8 << 30, num_streams: int = 2,
        planners: int = 8, timeout: float = 240.0, max_batch_configs: int = 1 << 20,
        launches_per_stream: int = 1) -> dict:
    if not os.path.exists(REPLAY):
        raise RuntimeError(f"{REPLAY} is not built: run `python -m rpmplv2_amd.build`")
    w = W.workload4(duration_s, rate, pool=1 << 20)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "serving.bin")
        write_problem(w, path, memory_budget, num_streams, max_batch_configs, planners, launches_per_stream)
        r = subprocess.run([REPLAY, path], capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"rvs_serve_replay failed ({r.returncode}): {r.stderr.strip()}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--duration", type=float, default=5.0)
    ap.add_argument("--rate", type=float, default=20000.0)
    ap.add_argument("--streams", type=int, default=2)
    ap.add_argument("--planners", type=int, default=8)
    ap.add_argument("--max-batch", type=int, default=1 << 20)
    ap.add_argument("--per-stream", type=int, default=1)
    args = ap.parse_args()
    print(json.dumps(run(args.duration, args.rate, num_streams=args.streams, planners=args.planners,
                         max_batch_configs=args.max_batch, launches_per_stream=args.per_stream)))


if __name__ == "__main__":
    main()
