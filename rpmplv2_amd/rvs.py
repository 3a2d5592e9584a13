"""Python binding of the engine's C ABI (include/rvs.h) over ctypes.

The engine itself is librvs.so (hand-written gfx950 kernels + C++ host code).
This module only marshals arguments: numpy arrays are passed as host memory,
torch CUDA tensors as device pointers (RVS_F_DEVICE).  There is no CPU
fallback: when librvs.so is missing or no GPU is present the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import robots as R
from . import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librvs.so")

RVS_OK, RVS_EINVAL, RVS_ENOMEM, RVS_EDEVICE, RVS_ERANGE, RVS_ENOTSUP = range(6)
PREC_REF32, PREC_FP64, PREC_FP32_REFINE = 0, 1, 2
F_DEVICE = 0x1
F_TIMING = 0x2
F_EAGER = 0x4
ENV_AABB, ENV_OBB = 0, 1
BOX_GROUND, BOX_DYNAMIC = 0x1, 0x2


def F_PREC(p: int) -> int:
    return (p + 1) << 4


class RvsError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"rvs status {status}: {msg}")
        self.status = status


class Options(C.Structure):
    _fields_ = [("device", C.c_int32), ("precision", C.c_int32), ("equality_threshold", C.c_float),
                ("max_distance", C.c_float), ("num_checks", C.c_int32), ("band", C.c_float)]


class Joint(C.Structure):
    _fields_ = [("type", C.c_int32), ("reserved", C.c_int32), ("xyz", C.c_double * 3), ("rpy", C.c_double * 3),
                ("axis", C.c_double * 3)]


class RobotDesc(C.Structure):
    _fields_ = [("skeleton_kind", C.c_int32), ("num_joints", C.c_int32), ("joints", C.POINTER(Joint)),
                ("num_links", C.c_int32), ("capsule_radius", C.POINTER(C.c_float)),
                ("ground_included", C.c_int32), ("gripper_length", C.c_float),
                ("num_self_pairs", C.c_int32), ("self_pairs", C.POINTER(C.c_int32))]


class Box(C.Structure):
    _fields_ = [("dim", C.c_float * 3), ("pos", C.c_float * 3), ("quat_wxyz", C.c_float * 4),
                ("flags", C.c_uint32), ("min_dist_tol", C.c_float)]


class EnvMotion(C.Structure):
    _fields_ = [("motion_type", C.c_int32), ("ws_center", C.c_float * 3), ("ws_radius", C.c_float),
                ("base_radius", C.c_float), ("robot_max_vel", C.c_float)]


MOTION_STRAIGHT, MOTION_CIRCULAR, MOTION_TWO_TUNNELS, MOTION_LIGHT_DIRECTIONS = range(4)


class ServerOptions(C.Structure):
    _fields_ = [("memory_budget", C.c_int64), ("num_streams", C.c_int32), ("max_batch_configs", C.c_int32),
                ("tick_seconds", C.c_float), ("precision", C.c_int32), ("launches_per_stream", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("robot", C.c_int32), ("env", C.c_int32), ("kind", C.c_int32), ("reserved", C.c_int32),
                ("n", C.c_int64), ("q", C.c_void_p), ("d", C.c_void_p), ("profile", C.c_void_p),
                ("flag", C.c_void_p)]


class RequestInfo(C.Structure):
    _fields_ = [("submit_s", C.c_double), ("admit_s", C.c_double), ("done_s", C.c_double), ("tick", C.c_int64),
                ("batch_requests", C.c_int32), ("status", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("configs_completed", C.c_int64),
                ("batches", C.c_int64), ("ticks", C.c_int64), ("max_inflight_batches", C.c_int32),
                ("reserved", C.c_int32), ("peak_bytes", C.c_int64)]


class Tuning(C.Structure):
    _fields_ = [("variant", C.c_int32), ("block", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("coherence_bits", C.c_int32), ("ms", C.c_float)]


class EdgeTuning(C.Structure):
    _fields_ = [("waves_per_cu", C.c_int32), ("slots_per_lane", C.c_int32), ("ms", C.c_float)]


REQ_VERDICT = 0
REQ_UNDERESTIMATION = 1
RVS_EBUSY = 6

EXPORTS = ["rvs_default_options", "rvs_create", "rvs_destroy", "rvs_last_error", "rvs_stream", "rvs_last_refined",
           "rvs_version", "rvs_robot_load", "rvs_env_create", "rvs_env_update", "rvs_env_bounds",
           "rvs_env_set_motion_law", "rvs_env_step", "rvs_distance",
           "rvs_distance_f64", "rvs_is_valid", "rvs_distance_async", "rvs_distance_async_static", "rvs_is_valid_async",
           "rvs_distance_underestimation_async", "rvs_self_collision_async",
           "rvs_distance_underestimation", "rvs_distance_underestimation_moving",
           "rvs_trajectory_is_safe", "rvs_self_distance",
           "rvs_self_collision", "rvs_enclosing_radii", "rvs_extend_spines", "rvs_extend_gen_spines",
           "rvs_edge_self_collision", "rvs_edge_is_valid", "rvs_edge_first_collision", "rvs_last_timing",
           "rvs_autotune", "rvs_set_tuning", "rvs_get_tuning", "rvs_set_small_batch", "rvs_autotune_edges", "rvs_set_edge_tuning", "rvs_server_default_options", "rvs_server_create", "rvs_server_destroy", "rvs_server_set_motion",
           "rvs_server_submit", "rvs_server_wait", "rvs_server_get_stats", "rvs_server_now",
           "rvs_server_last_error", "rvs_host_alloc", "rvs_host_free"]

_lib = None


def _stream_handle(stream):
    """hipStream_t of a torch.cuda.Stream (None stays None)."""
    return None if stream is None else C.c_void_p(stream.cuda_stream)


def _event_handle(event):
    """hipEvent_t of a torch.cuda.Event, created by a first record if torch
    has not created it yet (None stays None)."""
    if event is None:
        return None
    if not event.cuda_event:
        event.record()
    return C.c_void_p(event.cuda_event)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m rpmplv2_amd.build`")
        # torch ships its own HIP runtime under the same soname: loaded first,
        # it is the one librvs binds to, so that the process holds a single
        # runtime.  Loaded after librvs, it would start a second runtime that
        # finds no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, u32 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32
        ctx = C.c_void_p
        L.rvs_default_options.argtypes = [C.POINTER(Options)]
        L.rvs_default_options.restype = None
        L.rvs_create.argtypes = [C.POINTER(Options), C.POINTER(ctx)]
        L.rvs_destroy.argtypes = [ctx]
        L.rvs_destroy.restype = None
        L.rvs_last_error.argtypes = [ctx]
        L.rvs_last_error.restype = C.c_char_p
        L.rvs_stream.argtypes = [ctx]
        L.rvs_stream.restype = vp
        L.rvs_last_refined.argtypes = [ctx]
        L.rvs_last_refined.restype = i64
        L.rvs_version.restype = C.c_char_p
        L.rvs_last_timing.argtypes = [ctx, vp, i32]
        L.rvs_last_timing.restype = i32
        L.rvs_robot_load.argtypes = [ctx, C.POINTER(RobotDesc), C.POINTER(i32)]
        L.rvs_env_create.argtypes = [ctx, C.POINTER(Box), i32, i32, C.POINTER(i32)]
        L.rvs_env_update.argtypes = [ctx, i32, vp, vp, i32, u32]
        L.rvs_env_bounds.argtypes = [ctx, i32, vp, i32]
        L.rvs_env_set_motion_law.argtypes = [ctx, i32, C.POINTER(EnvMotion), vp, vp]
        L.rvs_env_step.argtypes = [ctx, i32, C.c_float, i32, vp, vp]
        L.rvs_distance.argtypes = [ctx, i32, i32, vp, i64, vp, vp, vp, vp, u32]
        L.rvs_distance_f64.argtypes = [ctx, i32, i32, vp, i64, vp, vp, vp, u32]
        L.rvs_is_valid.argtypes = [ctx, i32, i32, vp, i64, vp, u32]
        L.rvs_distance_async.argtypes = [ctx, i32, i32, vp, i64, vp, vp, vp, u32, vp, vp]
        L.rvs_distance_async_static.argtypes = [ctx, i32, i32, vp, i64, vp, vp, vp, u32, vp, vp]
        L.rvs_is_valid_async.argtypes = [ctx, i32, i32, vp, i64, vp, u32, vp, vp]
        L.rvs_distance_underestimation_async.argtypes = [ctx, i32, i32, vp, i64, vp, i64, vp, vp, vp, u32, vp, vp]
        L.rvs_self_collision_async.argtypes = [ctx, i32, vp, i64, vp, vp, u32, vp, vp]
        L.rvs_distance_underestimation.argtypes = [ctx, i32, i32, vp, i64, vp, i64, vp, vp, vp, u32]
        L.rvs_distance_underestimation_moving.argtypes = [ctx, i32, i32, vp, i64, vp, i64, vp, C.c_float, vp, vp, vp,
                                                          u32]
        L.rvs_trajectory_is_safe.argtypes = [ctx, i32, i32, vp, i64, i32, vp, vp, vp, vp, C.c_float, C.c_float, i32,
                                             vp, u32]
        L.rvs_self_distance.argtypes = [ctx, i32, vp, i64, vp, u32]
        L.rvs_self_collision.argtypes = [ctx, i32, vp, i64, vp, vp, u32]
        L.rvs_enclosing_radii.argtypes = [ctx, i32, vp, i64, vp, u32]
        L.rvs_edge_self_collision.argtypes = [ctx, i32, vp, vp, i64, C.c_float, vp, vp, u32]
        L.rvs_extend_spines.argtypes = [ctx, i32, vp, vp, vp, vp, i64, i32, i32, vp, vp, u32]
        L.rvs_extend_gen_spines.argtypes = [ctx, i32, i32, vp, vp, i64, i32, i32, i32, i32, C.c_float, vp, vp, vp,
                                            u32]
        L.rvs_autotune.argtypes = [ctx, i32, i32, vp, i64, u32, C.POINTER(Tuning)]
        L.rvs_set_tuning.argtypes = [ctx, i32, i32, C.POINTER(Tuning)]
        L.rvs_autotune_edges.argtypes = [ctx, i32, vp, vp, vp, vp, i64, i32, C.c_float, u32, C.POINTER(EdgeTuning)]
        L.rvs_set_edge_tuning.argtypes = [ctx, C.POINTER(EdgeTuning)]
        L.rvs_get_tuning.argtypes = [ctx, i32, i32, C.POINTER(Tuning)]
        L.rvs_set_small_batch.argtypes = [ctx, i64]
        L.rvs_server_default_options.argtypes = [C.POINTER(ServerOptions)]
        L.rvs_server_default_options.restype = None
        L.rvs_server_create.argtypes = [ctx, C.POINTER(ServerOptions), C.POINTER(C.c_void_p)]
        L.rvs_server_destroy.argtypes = [C.c_void_p]
        L.rvs_server_destroy.restype = None
        L.rvs_server_set_motion.argtypes = [C.c_void_p, i32, vp]
        L.rvs_server_submit.argtypes = [C.c_void_p, C.POINTER(Request), C.POINTER(C.c_uint64)]
        L.rvs_server_wait.argtypes = [C.c_void_p, C.c_uint64, i32, C.POINTER(RequestInfo)]
        L.rvs_server_get_stats.argtypes = [C.c_void_p, C.POINTER(ServerStats)]
        L.rvs_server_now.argtypes = [C.c_void_p]
        L.rvs_server_now.restype = C.c_double
        L.rvs_server_last_error.argtypes = [C.c_void_p]
        L.rvs_server_last_error.restype = C.c_char_p
        L.rvs_host_alloc.argtypes = [i64]
        L.rvs_host_alloc.restype = C.c_void_p
        L.rvs_host_free.argtypes = [C.c_void_p]
        L.rvs_host_free.restype = None
        L.rvs_edge_is_valid.argtypes = [ctx, i32, i32, vp, vp, i64, vp, u32]
        L.rvs_edge_first_collision.argtypes = [ctx, i32, vp, vp, vp, vp, i64, i32, C.c_float, vp, u32]
        for name in EXPORTS:
            f = getattr(L, name)
            if f.restype is C.c_int:
                f.restype = C.c_int
        _lib = L
    return _lib


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch") and hasattr(x, "data_ptr")


def _ptr(x):
    if x is None:
        return None
    if _is_torch(x):
        return C.c_void_p(x.data_ptr())
    return C.c_void_p(x.ctypes.data)


def robot_desc(rb: R.Robot):
    """Build an rvs_robot_desc; returns (desc, keepalive)."""
    js = (Joint * len(rb.joints))()
    for k, j in enumerate(rb.joints):
        js[k].type = j.type
        for i in range(3):
            js[k].xyz[i] = j.xyz[i]
            js[k].rpy[i] = j.rpy[i]
            js[k].axis[i] = j.axis[i]
    rad = np.ascontiguousarray(rb.radius, np.float32)
    pairs = np.ascontiguousarray(np.asarray(rb.self_pairs, np.int32).reshape(-1, 2))
    d = RobotDesc(rb.skeleton_kind, len(rb.joints), js, rb.num_links, rad.ctypes.data_as(C.POINTER(C.c_float)),
                  rb.ground_included, rb.gripper_length, len(pairs), pairs.ctypes.data_as(C.POINTER(C.c_int32)))
    return d, (js, rad, pairs)


def boxes_of(scene: W.Scene):
    arr = (Box * scene.num_obj)()
    for j in range(scene.num_obj):
        for i in range(3):
            arr[j].dim[i] = scene.dim[j, i]
            arr[j].pos[i] = scene.pos[j, i]
        for i in range(4):
            arr[j].quat_wxyz[i] = scene.quat[j, i]
        arr[j].flags = (BOX_GROUND if scene.is_ground[j] else 0) | \
            (BOX_DYNAMIC if scene.dynamic is not None and scene.dynamic[j] else 0)
        arr[j].min_dist_tol = scene.min_dist_tol[j]
    return arr


class Engine:
    """One engine context on one GPU (wraps rvs_context)."""

    def __init__(self, device: int = 0, precision: int = PREC_REF32, equality_threshold: float = 1e-4,
                 max_distance: float = 1000.0, num_checks: int = 10, band: float = 1e-4):
        L = lib()
        o = Options()
        L.rvs_default_options(C.byref(o))
        o.device, o.precision = device, precision
        o.equality_threshold, o.max_distance, o.num_checks, o.band = equality_threshold, max_distance, num_checks, band
        self._ctx = C.c_void_p()
        self._check(L.rvs_create(C.byref(o), C.byref(self._ctx)), "rvs_create")
        self.robots = {}
        self.envs = {}

    def close(self):
        if getattr(self, "_ctx", None):
            lib().rvs_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what):
        if st != RVS_OK:
            msg = lib().rvs_last_error(self._ctx).decode() if getattr(self, "_ctx", None) else what
            raise RvsError(st, f"{what}: {msg}")

    @property
    def stream(self) -> int:
        return lib().rvs_stream(self._ctx)

    def last_refined(self) -> int:
        return lib().rvs_last_refined(self._ctx)

    def last_timing(self):
        """Stage durations (ms) of the last call made with timing=True."""
        buf = np.zeros(4, np.float32)
        k = lib().rvs_last_timing(self._ctx, _ptr(buf), 4)
        return [float(x) for x in buf[:k]]

    # ---- set-up
    def load_robot(self, rb: R.Robot) -> int:
        d, keep = robot_desc(rb)
        rid = C.c_int32()
        self._check(lib().rvs_robot_load(self._ctx, C.byref(d), C.byref(rid)), "rvs_robot_load")
        self.robots[rid.value] = rb
        return rid.value

    def create_env(self, scene: W.Scene, semantics: int = ENV_AABB) -> int:
        arr = boxes_of(scene)
        eid = C.c_int32()
        self._check(lib().rvs_env_create(self._ctx, arr, scene.num_obj, semantics, C.byref(eid)), "rvs_env_create")
        self.envs[eid.value] = scene
        return eid.value

    def update_env(self, env: int, idx, pos):
        idx = np.ascontiguousarray(idx, np.int32)
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        self._check(lib().rvs_env_update(self._ctx, env, _ptr(idx), _ptr(pos), len(idx), 0), "rvs_env_update")

    def env_bounds(self, env: int) -> np.ndarray:
        m = self.envs[env].num_obj
        out = np.zeros((m, 6), np.float32)
        self._check(lib().rvs_env_bounds(self._ctx, env, _ptr(out), m), "rvs_env_bounds")
        return out

    # ---- queries on host arrays
    def distance(self, robot: int, env: int, q, precision: Optional[int] = None, want_nearest=False):
        rb = self.robots[robot]
        q = np.ascontiguousarray(q, np.float32).reshape(-1, rb.num_dofs)
        n, L, M = len(q), rb.num_links, self.envs[env].num_obj
        d = np.zeros(n, np.float32)
        prof = np.zeros((n, L), np.float32)
        ver = np.zeros(n, np.uint8)
        npo = np.zeros((n, M, L, 6), np.float32) if want_nearest else None
        fl = F_PREC(precision) if precision is not None else 0
        self._check(lib().rvs_distance(self._ctx, robot, env, _ptr(q), n, _ptr(d), _ptr(prof), _ptr(npo), _ptr(ver),
                                       fl), "rvs_distance")
        return d, prof, ver, npo

    def distance_f64(self, robot: int, env: int, q):
        rb = self.robots[robot]
        q = np.ascontiguousarray(q, np.float32).reshape(-1, rb.num_dofs)
        n, L = len(q), rb.num_links
        d = np.zeros(n, np.float64)
        prof = np.zeros((n, L), np.float64)
        ver = np.zeros(n, np.uint8)
        self._check(lib().rvs_distance_f64(self._ctx, robot, env, _ptr(q), n, _ptr(d), _ptr(prof), _ptr(ver), 0),
                    "rvs_distance_f64")
        return d, prof, ver

    def is_valid(self, robot: int, env: int, q, precision: Optional[int] = None, out=None, timing: bool = False):
        """numpy rows (returns the flags), or a torch device tensor with a
        uint8 device tensor out (RVS_F_DEVICE)."""
        rb = self.robots[robot]
        fl = (F_PREC(precision) if precision is not None else 0) | (F_TIMING if timing else 0)
        if _is_torch(q):
            self._check(lib().rvs_is_valid(self._ctx, robot, env, _ptr(q), q.shape[0], _ptr(out), fl | F_DEVICE),
                        "rvs_is_valid")
            return out
        q = np.ascontiguousarray(q, np.float32).reshape(-1, rb.num_dofs)
        out = np.zeros(len(q), np.uint8)
        self._check(lib().rvs_is_valid(self._ctx, robot, env, _ptr(q), len(q), _ptr(out), fl), "rvs_is_valid")
        return out

    def underestimation_device(self, robot: int, env: int, q, nearest_ref, d, profile, stored, timing: bool = False):
        """Device tensors throughout (RVS_F_DEVICE); nearest_ref [M][L][6]
        shared by all configurations or [n][M][L][6]."""
        rb = self.robots[robot]
        stride = 0 if nearest_ref.dim() == 3 else self.envs[env].num_obj * rb.num_links * 6
        fl = F_DEVICE | (F_TIMING if timing else 0)
        self._check(lib().rvs_distance_underestimation(self._ctx, robot, env, _ptr(q), q.shape[0], _ptr(nearest_ref),
                                                       stride, _ptr(d), _ptr(profile), _ptr(stored), fl),
                    "rvs_distance_underestimation")

    def underestimation(self, robot: int, env: int, q, nearest_ref):
        rb = self.robots[robot]
        q = np.ascontiguousarray(q, np.float32).reshape(-1, rb.num_dofs)
        n, L, M = len(q), rb.num_links, self.envs[env].num_obj
        npr = np.ascontiguousarray(nearest_ref, np.float32)
        stride = 0 if npr.ndim == 3 else M * L * 6
        d = np.zeros(n, np.float32)
        prof = np.zeros((n, L), np.float32)
        st = np.zeros(n, np.uint8)
        self._check(lib().rvs_distance_underestimation(self._ctx, robot, env, _ptr(q), n, _ptr(npr), stride, _ptr(d),
                                                       _ptr(prof), _ptr(st), 0), "rvs_distance_underestimation")
        return d, prof, st

    def self_distance(self, robot: int, q):
        rb = self.robots[robot]
        q = np.ascontiguousarray(q, np.float32).reshape(-1, rb.num_dofs)
        out = np.zeros((len(q), len(rb.self_pairs)), np.float32)
        self._check(lib().rvs_self_distance(self._ctx, robot, _ptr(q), len(q), _ptr(out), 0), "rvs_self_distance")
        return out

    def self_collision(self, robot: int, q):
        """(min pair distance [N], collision-free flag [N]) of the robot's self pairs."""
        rb = self.robots[robot]
        q = np.ascontiguousarray(q, np.float32).reshape(-1, rb.num_dofs)
        m = np.zeros(len(q), np.float32)
        f = np.zeros(len(q), np.uint8)
        self._check(lib().rvs_self_collision(self._ctx, robot, _ptr(q), len(q), _ptr(m), _ptr(f), 0),
                    "rvs_self_collision")
        return m, f

    def enclosing_radii(self, robot: int, q):
        """AbstractRobot::computeEnclosingRadii: (N, dofs, dofs + 1)."""
        rb = self.robots[robot]
        q = np.ascontiguousarray(q, np.float32).reshape(-1, rb.num_dofs)
        D = rb.num_dofs
        out = np.zeros((len(q), D, D + 1), np.float32)
        self._check(lib().rvs_enclosing_radii(self._ctx, robot, _ptr(q), len(q), _ptr(out), 0), "rvs_enclosing_radii")
        return out

    def set_small_batch(self, max_configs: int):
        """Reference-arithmetic batches up to max_configs run one wave per configuration."""
        self._check(lib().rvs_set_small_batch(self._ctx, int(max_configs)), "rvs_set_small_batch")

    def set_motion_law(self, env: int, motion_type: int, vel, max_vel, ws_center=(0.0, 0.0, 0.0),
                       ws_radius: float = 1.0, base_radius: float = 0.1, robot_max_vel: float = 1.0):
        """Motion law of Environment::updateEnvironment for an environment."""
        m = EnvMotion(motion_type, (C.c_float * 3)(*ws_center), ws_radius, base_radius, robot_max_vel)
        v = np.ascontiguousarray(vel, np.float32)
        mv = np.ascontiguousarray(max_vel, np.float32)
        self._check(lib().rvs_env_set_motion_law(self._ctx, env, C.byref(m), _ptr(v), _ptr(mv)),
                    "rvs_env_set_motion_law")

    def env_step(self, env: int, delta_time: float, steps: int = 1):
        """Environment::updateEnvironment(delta_time), steps times: (pos, vel) (M, 3)."""
        M = self.envs[env].num_obj
        pos = np.zeros((M, 3), np.float32)
        vel = np.zeros((M, 3), np.float32)
        self._check(lib().rvs_env_step(self._ctx, env, float(delta_time), int(steps), _ptr(pos), _ptr(vel)),
                    "rvs_env_step")
        return pos, vel

    def underestimation_moving(self, robot: int, env: int, q, nearest_ref, max_vel, delta_t: float):
        """AbstractTrajectory::computeDistanceUnderestimation with planes
        moved for delta_t: (d, profile, stored)."""
        rb = self.robots[robot]
        q = np.ascontiguousarray(q, np.float32).reshape(-1, rb.num_dofs)
        n, L = len(q), rb.num_links
        npr = np.ascontiguousarray(nearest_ref, np.float32)
        stride = 0 if npr.ndim == 3 else npr[0].size
        mv = np.ascontiguousarray(max_vel, np.float32)
        d = np.zeros(n, np.float32)
        prof = np.zeros((n, L), np.float32)
        st = np.zeros(n, np.uint8)
        self._check(lib().rvs_distance_underestimation_moving(self._ctx, robot, env, _ptr(q), n, _ptr(npr), stride,
                                                              _ptr(mv), float(delta_t), _ptr(d), _ptr(prof), _ptr(st),
                                                              0), "rvs_distance_underestimation_moving")
        return d, prof, st

    def trajectory_is_safe(self, robot: int, env: int, points, q_current, profile_current, nearest_current, max_vel,
                           t_iter: float = 0.0, time_step: float = 0.01, self_collision: bool = False):
        """AbstractTrajectory::isSafe for trajectories points (T, K, dofs)
        leaving q_current: safe (T,) uint8."""
        rb = self.robots[robot]
        pts = np.ascontiguousarray(points, np.float32)
        T, K = pts.shape[0], pts.shape[1]
        args = [np.ascontiguousarray(x, np.float32) for x in (q_current, profile_current, nearest_current, max_vel)]
        safe = np.zeros(T, np.uint8)
        self._check(lib().rvs_trajectory_is_safe(self._ctx, robot, env, _ptr(pts), T, K, *[_ptr(x) for x in args],
                                                 float(t_iter), float(time_step), 1 if self_collision else 0,
                                                 _ptr(safe), 0), "rvs_trajectory_is_safe")
        return safe

    def edge_self_collision(self, robot: int, q1, q2, eps_step: float = 0.1):
        """Capsule part of xArm6::checkSelfCollision(q1, q2) per edge:
        (collide (E,) uint8, q2_out (E, dofs))."""
        rb = self.robots[robot]
        q1 = np.ascontiguousarray(q1, np.float32).reshape(-1, rb.num_dofs)
        q2 = np.ascontiguousarray(q2, np.float32).reshape(q1.shape)
        col = np.zeros(len(q1), np.uint8)
        out = np.zeros_like(q1)
        self._check(lib().rvs_edge_self_collision(self._ctx, robot, _ptr(q1), _ptr(q2), len(q1), float(eps_step),
                                                  _ptr(col), _ptr(out), 0), "rvs_edge_self_collision")
        return col, out

    def extend_spines(self, robot: int, q, d, profile, q_e, num_iter: int = 5, expanded: bool = True):
        """RBTConnect::extendSpine for a batch of spines (robot without
        self-collision checking): (q_new (N, dofs), status (N,) int32:
        1 advanced, 2 trapped, 3 reached)."""
        rb = self.robots[robot]
        D, L = rb.num_dofs, rb.num_links
        q = np.ascontiguousarray(q, np.float32).reshape(-1, D)
        n = len(q)
        d = np.ascontiguousarray(d, np.float32).reshape(n)
        prof = None if profile is None else np.ascontiguousarray(profile, np.float32).reshape(n, L)
        qe = np.ascontiguousarray(q_e, np.float32).reshape(n, D)
        qn = np.zeros((n, D), np.float32)
        st = np.zeros(n, np.int32)
        self._check(lib().rvs_extend_spines(self._ctx, robot, _ptr(q), _ptr(d), _ptr(prof), _ptr(qe), n, int(num_iter),
                                            1 if expanded else 0, _ptr(qn), _ptr(st), 0), "rvs_extend_spines")
        return qn, st

    def extend_gen_spines(self, robot: int, env: int, q_root, q_e, num_layers: int = 5, num_iter: int = 5,
                          expanded: bool = True, use_real_dist: bool = False, d_crit: float = 0.03):
        """RGBTConnect::extendGenSpine from one root towards each target:
        (q_new (N, dofs), status (N,), d (N,))."""
        rb = self.robots[robot]
        D = rb.num_dofs
        qr = np.ascontiguousarray(q_root, np.float32).reshape(D)
        qe = np.ascontiguousarray(q_e, np.float32).reshape(-1, D)
        n = len(qe)
        qn = np.zeros((n, D), np.float32)
        st = np.zeros(n, np.int32)
        d = np.zeros(n, np.float32)
        self._check(lib().rvs_extend_gen_spines(self._ctx, robot, env, _ptr(qr), _ptr(qe), n, int(num_layers),
                                                int(num_iter), 1 if expanded else 0, 1 if use_real_dist else 0,
                                                float(d_crit), _ptr(qn), _ptr(st), _ptr(d), 0),
                    "rvs_extend_gen_spines")
        return qn, st, d

    def extend_gen_spines_device(self, robot: int, env: int, q_root, q_e, q_new, status, d, num_layers: int = 5,
                                 num_iter: int = 5, expanded: bool = True, use_real_dist: bool = False,
                                 d_crit: float = 0.03):
        """Device-tensor form of extend_gen_spines (torch CUDA tensors)."""
        self._check(lib().rvs_extend_gen_spines(self._ctx, robot, env, _ptr(q_root), _ptr(q_e), q_e.shape[0],
                                                int(num_layers), int(num_iter), 1 if expanded else 0,
                                                1 if use_real_dist else 0, float(d_crit), _ptr(q_new), _ptr(status),
                                                _ptr(d), F_DEVICE), "rvs_extend_gen_spines")

    def self_collision_device(self, robot: int, q, min_dist=None, free_flag=None):
        self._check(lib().rvs_self_collision(self._ctx, robot, _ptr(q), q.shape[0], _ptr(min_dist), _ptr(free_flag),
                                             F_DEVICE), "rvs_self_collision")

    # ---- queries on device tensors (torch CUDA tensors; no staging)
    def distance_device(self, robot: int, env: int, q, d, profile=None, verdict=None, nearest=None,
                        precision: Optional[int] = None, timing: bool = False, eager: bool = False):
        fl = (F_DEVICE | (F_PREC(precision) if precision is not None else 0) | (F_TIMING if timing else 0)
              | (F_EAGER if eager else 0))
        self._check(lib().rvs_distance(self._ctx, robot, env, _ptr(q), q.shape[0], _ptr(d), _ptr(profile),
                                       _ptr(nearest), _ptr(verdict), fl), "rvs_distance")

    def distance_async(self, robot: int, env: int, q, d, profile=None, verdict=None, precision: Optional[int] = None,
                       stream=None, done=None):
        """rvs_distance_async on device tensors: enqueued on `stream` (a
        torch.cuda.Stream; None = the context's stream), returns at once;
        `done` (a torch.cuda.Event, may be None) is recorded after the work."""
        fl = F_DEVICE | (F_PREC(precision) if precision is not None else 0)
        self._check(lib().rvs_distance_async(self._ctx, robot, env, _ptr(q), q.shape[0], _ptr(d), _ptr(profile),
                                             _ptr(verdict), fl, _stream_handle(stream), _event_handle(done)),
                    "rvs_distance_async")

    def distance_async_static(self, robot: int, env: int, q, d, profile=None, verdict=None,
                              precision: Optional[int] = None, stream=None, done=None):
        """rvs_distance_async_static on device tensors: rvs_distance_async
        with its scratch in the library's fixed device array (no allocation;
        at most 1,048,576 configurations)."""
        fl = F_DEVICE | (F_PREC(precision) if precision is not None else 0)
        self._check(lib().rvs_distance_async_static(self._ctx, robot, env, _ptr(q), q.shape[0], _ptr(d),
                                                    _ptr(profile), _ptr(verdict), fl, _stream_handle(stream),
                                                    _event_handle(done)), "rvs_distance_async_static")

    def is_valid_async(self, robot: int, env: int, q, valid, precision: Optional[int] = None, stream=None,
                       done=None):
        """rvs_is_valid_async on device tensors (see distance_async)."""
        fl = F_DEVICE | (F_PREC(precision) if precision is not None else 0)
        self._check(lib().rvs_is_valid_async(self._ctx, robot, env, _ptr(q), q.shape[0], _ptr(valid), fl,
                                             _stream_handle(stream), _event_handle(done)), "rvs_is_valid_async")

    def underestimation_async(self, robot: int, env: int, q, nearest_ref, d, profile=None, stored=None,
                              per_config: bool = False, stream=None, done=None):
        """rvs_distance_underestimation_async on device tensors (see
        distance_async); per_config: one nearest-point set per configuration."""
        rb, M = self.robots[robot], self.envs[env].num_obj
        stride = M * rb.num_links * 6 if per_config else 0
        self._check(lib().rvs_distance_underestimation_async(
            self._ctx, robot, env, _ptr(q), q.shape[0], _ptr(nearest_ref), stride, _ptr(d), _ptr(profile),
            _ptr(stored), F_DEVICE, _stream_handle(stream), _event_handle(done)), "rvs_distance_underestimation_async")

    def self_collision_async(self, robot: int, q, min_dist=None, free=None, stream=None, done=None):
        """rvs_self_collision_async on device tensors (see distance_async)."""
        self._check(lib().rvs_self_collision_async(self._ctx, robot, _ptr(q), q.shape[0], _ptr(min_dist), _ptr(free),
                                                   F_DEVICE, _stream_handle(stream), _event_handle(done)),
                    "rvs_self_collision_async")

    def distance_f64_device(self, robot: int, env: int, q, d, profile=None, verdict=None, timing: bool = False):
        fl = F_DEVICE | (F_TIMING if timing else 0)
        self._check(lib().rvs_distance_f64(self._ctx, robot, env, _ptr(q), q.shape[0], _ptr(d), _ptr(profile),
                                           _ptr(verdict), fl), "rvs_distance_f64")

    # ---- autotuning of the FP32 bulk pass
    def autotune(self, robot: int, env: int, q) -> dict:
        t = Tuning()
        dev = _is_torch(q)
        if not dev:
            q = np.ascontiguousarray(q, np.float32).reshape(-1, self.robots[robot].num_dofs)
        self._check(lib().rvs_autotune(self._ctx, robot, env, _ptr(q), q.shape[0], F_DEVICE if dev else 0,
                                       C.byref(t)), "rvs_autotune")
        return {"variant": t.variant, "block": t.block, "blocks_per_cu": t.blocks_per_cu,
                "coherence_bits": t.coherence_bits, "ms": t.ms}

    def set_tuning(self, robot: int, env: int, variant: int, block: int = 256, blocks_per_cu: int = 0,
                   coherence_bits: int = 0):
        t = Tuning(variant, block, blocks_per_cu, coherence_bits, 0.0)
        self._check(lib().rvs_set_tuning(self._ctx, robot, env, C.byref(t)), "rvs_set_tuning")

    def get_tuning(self, robot: int, env: int) -> dict:
        t = Tuning()
        self._check(lib().rvs_get_tuning(self._ctx, robot, env, C.byref(t)), "rvs_get_tuning")
        return {"variant": t.variant, "block": t.block, "blocks_per_cu": t.blocks_per_cu,
                "coherence_bits": t.coherence_bits}

    # ---- edges
    def edge_is_valid(self, robot: int, env: int, q1, q2, out=None, timing: bool = False):
        """numpy rows (returns the flags), or torch device tensors with a
        uint8 device tensor out (RVS_F_DEVICE)."""
        rb = self.robots[robot]
        fl = F_TIMING if timing else 0
        if _is_torch(q1):
            self._check(lib().rvs_edge_is_valid(self._ctx, robot, env, _ptr(q1), _ptr(q2), q1.shape[0], _ptr(out),
                                                fl | F_DEVICE), "rvs_edge_is_valid")
            return out
        q1 = np.ascontiguousarray(q1, np.float32).reshape(-1, rb.num_dofs)
        q2 = np.ascontiguousarray(q2, np.float32).reshape(-1, rb.num_dofs)
        out = np.zeros(len(q1), np.uint8)
        self._check(lib().rvs_edge_is_valid(self._ctx, robot, env, _ptr(q1), _ptr(q2), len(q1), _ptr(out), fl),
                    "rvs_edge_is_valid")
        return out

    def edge_first_collision(self, env: int, robot_ids, q1, direction, K, step: float = 0.05):
        """Edges of several robots in one launch; q1/direction rows padded to
        the largest joint count.  Returns the first colliding state per edge."""
        robot_ids = np.ascontiguousarray(robot_ids, np.int32)
        q1 = np.ascontiguousarray(q1, np.float32)
        direction = np.ascontiguousarray(direction, np.float32)
        K = np.ascontiguousarray(K, np.int32)
        n = len(robot_ids)
        out = np.zeros(n, np.int32)
        self._check(lib().rvs_edge_first_collision(self._ctx, env, _ptr(robot_ids), _ptr(q1), _ptr(direction), _ptr(K),
                                                   n, q1.shape[1], C.c_float(step), _ptr(out), 0),
                    "rvs_edge_first_collision")
        return out

    def stack_edges(self, batches):
        """Stack EdgeBatch objects (one robot each; robots loaded once and
        cached) into (robot_ids, q1, dir, K) rows padded to the widest robot."""
        cache = self.__dict__.setdefault("_edge_robot_ids", {})
        ids, rows_q, rows_u, ks = [], [], [], []
        width = max(b.robot.num_dofs for b in batches)
        for b in batches:
            key = id(b.robot)
            if key not in cache:
                cache[key] = (self.load_robot(b.robot), b.robot)
            rid = cache[key][0]
            n = len(b.q1)
            pq = np.zeros((n, width), np.float32); pq[:, :b.robot.num_dofs] = b.q1
            pu = np.zeros((n, width), np.float32); pu[:, :b.robot.num_dofs] = b.dir
            ids.append(np.full(n, rid, np.int32)); rows_q.append(pq); rows_u.append(pu); ks.append(b.K)
        return np.concatenate(ids), np.concatenate(rows_q), np.concatenate(rows_u), np.concatenate(ks)

    def edge_first_collision_device(self, env: int, robot_ids, q1, direction, K, first, step: float = 0.05,
                                    timing: bool = False):
        """robot_ids and K host arrays; q1, direction and first device tensors."""
        robot_ids = np.ascontiguousarray(robot_ids, np.int32)
        K = np.ascontiguousarray(K, np.int32)
        fl = F_DEVICE | (F_TIMING if timing else 0)
        self._check(lib().rvs_edge_first_collision(self._ctx, env, _ptr(robot_ids), _ptr(q1), _ptr(direction), _ptr(K),
                                                   len(robot_ids), q1.shape[1], C.c_float(step), _ptr(first), fl),
                    "rvs_edge_first_collision")

    def autotune_edges(self, env: int, robot_ids, q1, direction, K, step: float = 0.05) -> dict:
        robot_ids = np.ascontiguousarray(robot_ids, np.int32)
        K = np.ascontiguousarray(K, np.int32)
        dev = _is_torch(q1)
        if not dev:
            q1 = np.ascontiguousarray(q1, np.float32)
            direction = np.ascontiguousarray(direction, np.float32)
        t = EdgeTuning()
        self._check(lib().rvs_autotune_edges(self._ctx, env, _ptr(robot_ids), _ptr(q1), _ptr(direction), _ptr(K),
                                             len(robot_ids), q1.shape[1], C.c_float(step), F_DEVICE if dev else 0,
                                             C.byref(t)), "rvs_autotune_edges")
        return {"waves_per_cu": t.waves_per_cu, "slots_per_lane": t.slots_per_lane, "ms": t.ms}

    def set_edge_tuning(self, waves_per_cu: int, slots_per_lane: int):
        t = EdgeTuning(waves_per_cu, slots_per_lane, 0.0)
        self._check(lib().rvs_set_edge_tuning(self._ctx, C.byref(t)), "rvs_set_edge_tuning")

    def edge_batches(self, env: int, batches):
        """Workload-3 helper: stack EdgeBatch objects (one robot each) into a
        single first-collision call; returns one array per batch."""
        ids, q1, u, K = self.stack_edges(batches)
        first = self.edge_first_collision(env, ids, q1, u, K, batches[0].step)
        out, o = [], 0
        for b in batches:
            out.append(first[o:o + len(b.q1)])
            o += len(b.q1)
        return out


class Server:
    """Request server over an Engine's context (rvs_server_*).  Results land
    in numpy arrays owned by this object until wait() returns them."""

    def __init__(self, engine: Engine, memory_budget: int = 8 << 30, num_streams: int = 2,
                 max_batch_configs: int = 1 << 20, tick_seconds: float = 0.01, precision: int = PREC_FP32_REFINE,
                 launches_per_stream: int = 1):
        L = lib()
        o = ServerOptions()
        L.rvs_server_default_options(C.byref(o))
        o.memory_budget, o.num_streams, o.max_batch_configs = memory_budget, num_streams, max_batch_configs
        o.launches_per_stream = launches_per_stream
        o.tick_seconds, o.precision = tick_seconds, precision
        self.engine = engine
        self._srv = C.c_void_p()
        st = L.rvs_server_create(engine._ctx, C.byref(o), C.byref(self._srv))
        if st != RVS_OK:
            raise RvsError(st, "rvs_server_create")
        self._live = {}

    def _check(self, st, what):
        if st != RVS_OK:
            raise RvsError(st, f"{what}: {lib().rvs_server_last_error(self._srv).decode()}")

    def set_motion(self, env: int, vel):
        v = np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
        self._check(lib().rvs_server_set_motion(self._srv, env, _ptr(v)), "rvs_server_set_motion")

    def now(self) -> float:
        return lib().rvs_server_now(self._srv)

    def submit(self, robot: int, env: int, kind: int, q, want_profile: bool = False) -> int:
        rb = self.engine.robots[robot]
        q = np.ascontiguousarray(q, np.float32).reshape(-1, rb.num_dofs)
        n = len(q)
        d = np.zeros(n, np.float32)
        flag = np.zeros(n, np.uint8)
        prof = np.zeros((n, rb.num_links), np.float32) if want_profile else None
        r = Request(robot, env, kind, 0, n, q.ctypes.data if n else None, d.ctypes.data if n else None,
                    prof.ctypes.data if (prof is not None and n) else None, flag.ctypes.data if n else None)
        t = C.c_uint64()
        self._check(lib().rvs_server_submit(self._srv, C.byref(r), C.byref(t)), "rvs_server_submit")
        self._live[t.value] = (q, d, flag, prof)
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1):
        """(d, flag, profile, info dict) once done; None on timeout."""
        info = RequestInfo()
        st = lib().rvs_server_wait(self._srv, ticket, timeout_ms, C.byref(info))
        if st == RVS_EBUSY:
            return None
        self._check(st, "rvs_server_wait")
        q, d, flag, prof = self._live.pop(ticket)
        return d, flag, prof, {"submit_s": info.submit_s, "admit_s": info.admit_s, "done_s": info.done_s,
                               "tick": info.tick, "batch_requests": info.batch_requests}

    def stats(self) -> dict:
        s = ServerStats()
        self._check(lib().rvs_server_get_stats(self._srv, C.byref(s)), "rvs_server_get_stats")
        return {k: getattr(s, k) for k, _ in ServerStats._fields_ if k != "reserved"}

    def close(self):
        if getattr(self, "_srv", None):
            lib().rvs_server_destroy(self._srv)
            self._srv = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
