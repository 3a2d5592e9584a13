"""Seeded synthetic workloads for the capsule collision / distance engine.

Every draw uses numpy's PCG64 generator seeded with 1000 + workload number,
in double, rounded to float32 where the reference stores float.  Scenes are
lists of boxes (dim, pos, quaternion w-x-y-z, ground flag, min_dist_tol);
their axis-aligned bounds follow FCL's rule for CollisionObject::computeAABB
(identity rotation: pos -+ dim/2; otherwise pos -+ |dim/2| on every axis),
restated in Box.cpp:3-15 / Object.cpp:21-26 of the reference.

Workloads (sizes as quoted; each function takes a scale-down factor for tests):
  1  planar 2-DOF, r = 0, 4 boxes in a 2 m disk, 65,536 configurations
  2  xArm6, r = 0.05, ground box + 16 boxes in a 1 m ball at (0,0,0.4),
     boxes that make the start pose (q = 0) collide rejected, 1,048,576
     configurations
  3  chains of 10/14/18/22 DOF with random link lengths, r = 0.04, self-collision
     on, 64 boxes in a 1.5 m ball, 4,096 discretised edges per chain
  4  serving stream (see serving.py)
  5  22-DOF chain, r = 0.05, 512 rotated boxes in a 2 m ball, 16,777,216 configurations
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import robots as R


@dataclass
class Scene:
    dim: np.ndarray          # (M, 3) float32
    pos: np.ndarray          # (M, 3) float32
    quat: np.ndarray         # (M, 4) float32, w x y z
    is_ground: np.ndarray    # (M,) uint8
    min_dist_tol: np.ndarray # (M,) float32
    aabb: np.ndarray = field(init=False)
    dynamic: Optional[np.ndarray] = None   # (M,) uint8, 1 = label "dynamic_obstacle"

    def __post_init__(self):
        self.dim = np.ascontiguousarray(self.dim, np.float32).reshape(-1, 3)
        self.pos = np.ascontiguousarray(self.pos, np.float32).reshape(-1, 3)
        self.quat = np.ascontiguousarray(self.quat, np.float32).reshape(-1, 4)
        self.is_ground = np.ascontiguousarray(self.is_ground, np.uint8)
        self.min_dist_tol = np.ascontiguousarray(self.min_dist_tol, np.float32)
        self.aabb = aabb_of(self.dim, self.pos, self.quat)

    @property
    def num_obj(self) -> int:
        return len(self.dim)

    def obb_table(self) -> np.ndarray:
        """[M][15]: centre, rotation rows, half extents - the oriented-box
        table the engine builds (unit quaternion -> matrix in double)."""
        q = self.quat.astype(np.float64)
        w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
        nrm = np.sqrt(((w * w + x * x) + y * y) + z * z)   # left to right, as the C ABI does
        w, x, y, z = w / nrm, x / nrm, y / nrm, z / nrm
        R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                      2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                      2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], 1)
        out = np.empty((self.num_obj, 15), np.float32)
        out[:, :3] = self.pos
        out[:, 3:12] = R.astype(np.float32)
        out[:, 12:] = np.float32(0.5) * self.dim
        return out

    def subset(self, idx) -> "Scene":
        idx = np.asarray(idx)
        return Scene(self.dim[idx], self.pos[idx], self.quat[idx], self.is_ground[idx], self.min_dist_tol[idx])

    def concat(self, other: "Scene") -> "Scene":
        return Scene(np.concatenate([self.dim, other.dim]), np.concatenate([self.pos, other.pos]),
                     np.concatenate([self.quat, other.quat]), np.concatenate([self.is_ground, other.is_ground]),
                     np.concatenate([self.min_dist_tol, other.min_dist_tol]))


MAX_DISTANCE = np.float32(1000.0)


def _quat_matrix_f32(q):
    """Eigen Quaternion<float>::toRotationMatrix, vectorised (float32)."""
    q = np.asarray(q, np.float32)
    w, x, y, z = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    two = np.float32(2)
    one = np.float32(1)
    tx, ty, tz = two * x, two * y, two * z
    twx, twy, twz = tx * w, ty * w, tz * w
    txx, txy, txz = tx * x, ty * x, tz * x
    tyy, tyz, tzz = ty * y, tz * y, tz * z
    return np.stack([np.stack([one - (tyy + tzz), txy - twz, txz + twy], -1),
                     np.stack([txy + twz, one - (txx + tzz), tyz - twx], -1),
                     np.stack([txz - twy, tyz + twx, one - (txx + tyy)], -1)], -2)


def is_identity_rot(Rm) -> np.ndarray:
    """Eigen isIdentity with float dummy precision 1e-5."""
    Rm = np.asarray(Rm, np.float32)
    prec = np.float32(1e-5)
    eye = np.eye(3, dtype=bool)
    diag_ok = np.abs(Rm - 1) <= prec * np.minimum(np.abs(Rm), 1)
    off_ok = np.abs(Rm) <= prec
    ok = np.where(eye, diag_ok, off_ok)
    return ok.all(axis=(-1, -2))


def aabb_of(dim, pos, quat) -> np.ndarray:
    dim = np.asarray(dim, np.float32).reshape(-1, 3)
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    quat = np.asarray(quat, np.float32).reshape(-1, 4)
    h = np.float32(0.5) * dim
    ident = is_identity_rot(_quat_matrix_f32(quat))
    out = np.empty((len(dim), 6), np.float32)
    out[:, :3] = -h + pos
    out[:, 3:] = h + pos
    if (~ident).any():
        m = -h[~ident]
        rad = np.sqrt(m[:, 0] * m[:, 0] + (m[:, 1] * m[:, 1] + m[:, 2] * m[:, 2])).astype(np.float32)
        p = pos[~ident]
        out[~ident, :3] = p - rad[:, None]
        out[~ident, 3:] = p + rad[:, None]
    return out


def ground_box() -> Scene:
    """The repository's ground box (data/xarm6/scenario_test/scenario_test.yaml:2-6)."""
    return Scene([[1.5, 1.5, 0.1]], [[0, 0, -0.05]], [[1, 0, 0, 0]], [1], [0.05])


def _uniform_ball(rng, n, radius, center=(0, 0, 0)):
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    r = radius * rng.random(n) ** (1.0 / 3.0)
    return np.asarray(center, np.float64) + u * r[:, None]


def _boxes(rng, n, side_lo, side_hi, radius, center=(0, 0, 0), rotated=False, tol=None) -> Scene:
    dim = rng.uniform(side_lo, side_hi, size=(n, 3))
    pos = _uniform_ball(rng, n, radius, center)
    if rotated:
        quat = _shoemake(rng, n)
    else:
        quat = np.tile([1.0, 0.0, 0.0, 0.0], (n, 1))
    t = np.full(n, MAX_DISTANCE if tol is None else tol, np.float32)
    return Scene(dim, pos, quat, np.zeros(n, np.uint8), t)


def _shoemake(rng, n):
    """Uniform random unit quaternions (Shoemake 1992), returned w x y z."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    x = a * np.sin(2 * np.pi * u2)
    y = a * np.cos(2 * np.pi * u2)
    z = b * np.sin(2 * np.pi * u3)
    w = b * np.cos(2 * np.pi * u3)
    return np.stack([w, x, y, z], 1)


def seg_box_distance(A, B, lo, hi) -> float:
    """Exact distance from segment AB to the box [lo, hi] (float64).  The
    squared distance of P(t) = A + t (B - A) to the box is convex and
    piecewise quadratic in t, its pieces split where a coordinate crosses a
    slab bound; the minimum is at a breakpoint or at a piece's vertex."""
    A = np.asarray(A, np.float64)
    B = np.asarray(B, np.float64)
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    D = B - A
    ts = [0.0, 1.0]
    for k in range(3):
        if D[k] != 0:
            for v in (lo[k], hi[k]):
                t = (v - A[k]) / D[k]
                if 0 < t < 1:
                    ts.append(t)
    ts.sort()

    def d2(t):
        e = np.maximum(np.maximum(lo - (A + t * D), (A + t * D) - hi), 0.0)
        return float(e @ e)

    best = min(d2(t) for t in ts)
    for a, b in zip(ts[:-1], ts[1:]):
        P = A + 0.5 * (a + b) * D
        num = den = 0.0
        for k in range(3):       # axes outside the slab on this piece: e_k = c_k + D_k t
            if P[k] < lo[k]:
                c = A[k] - lo[k]
            elif P[k] > hi[k]:
                c = A[k] - hi[k]
            else:
                continue
            num += c * D[k]
            den += D[k] * D[k]
        if den > 0:
            best = min(best, d2(min(max(-num / den, a), b)))
    return float(np.sqrt(best))


def collides_at_pose(rb: R.Robot, aabb: np.ndarray, q) -> bool:
    """Some capsule of the pose q lies closer to the box than its radius:
    the geometric form of isValid(q) failing for one box (the collision tree
    of RealVectorSpace.cpp:277-309 reports the same except in the corner
    regions of its index quirk, CollisionAndDistance.cpp:69-72)."""
    P = R.skeleton_np(rb, q)[0]
    return any(seg_box_distance(P[i], P[i + 1], aabb[:3], aabb[3:]) < rb.radius[i] for i in range(rb.num_links))


def uniform_configs(rng, n, limits) -> np.ndarray:
    lo = np.asarray(limits, np.float64)[:, 0]
    hi = np.asarray(limits, np.float64)[:, 1]
    return (lo + (hi - lo) * rng.random((n, len(lo)))).astype(np.float32)


# ---------------------------------------------------------------- workload 1
@dataclass
class ConfigBatch:
    name: str
    robot: R.Robot
    scene: Scene
    q: np.ndarray


def workload1(n: int = 65536) -> ConfigBatch:
    rng = np.random.default_rng(1001)
    rb = R.planar_2dof(radius=0.0)
    m = 4
    rr = 2.0 * np.sqrt(rng.random(m))
    th = 2 * np.pi * rng.random(m)
    pos = np.stack([rr * np.cos(th), rr * np.sin(th), np.zeros(m)], 1)
    dim = np.stack([rng.uniform(0.1, 0.5, m), rng.uniform(0.1, 0.5, m), np.full(m, 0.2)], 1)
    sc = Scene(dim, pos, np.tile([1.0, 0, 0, 0], (m, 1)), np.zeros(m, np.uint8), np.full(m, MAX_DISTANCE))
    q = (rng.uniform(-np.pi, np.pi, size=(n, 2))).astype(np.float32)
    return ConfigBatch("planar2dof_4aabb", rb, sc, q)


# ---------------------------------------------------------------- workload 2
def scene2(rng=None, num_boxes: int = 16) -> Scene:
    rng = rng or np.random.default_rng(1002)
    rb = R.xarm6()
    q0 = np.zeros(6)
    kept: List[Scene] = []
    while len(kept) < num_boxes:
        b = _boxes(rng, 1, 0.05, 0.3, 1.0, (0, 0, 0.4))
        if collides_at_pose(rb, b.aabb[0], q0):     # CommonFunctions.h:105-110: isValid(start) fails
            continue
        kept.append(b)
    boxes = kept[0]
    for b in kept[1:]:
        boxes = boxes.concat(b)
    return ground_box().concat(boxes)


def workload2(n: int = 1048576) -> ConfigBatch:
    rng = np.random.default_rng(1002)
    sc = scene2(rng)
    rb = R.xarm6(radius=0.05, ground_included=1)
    q = uniform_configs(rng, n, rb.limits)
    return ConfigBatch("xarm6_16aabb_ground", rb, sc, q)


# ---------------------------------------------------------------- workload 3
@dataclass
class EdgeBatch:
    robot: R.Robot
    q1: np.ndarray     # (E, n)
    dir: np.ndarray    # (E, n) unit vectors
    K: np.ndarray      # (E,) states per edge
    step: float = 0.05


@dataclass
class EdgeWorkload:
    scene: Scene
    batches: List[EdgeBatch]


def workload3(edges_per_robot: int = 4096, dofs=(10, 14, 18, 22)) -> EdgeWorkload:
    rng = np.random.default_rng(1003)
    sc = _boxes(rng, 64, 0.1, 0.5, 1.5)
    batches = []
    for n in dofs:
        lengths = rng.uniform(0.1, 0.3, n)
        rb = R.chain(lengths, radius=0.04, name=f"chain{n}", self_collision=True)
        q1 = rng.uniform(-np.pi, np.pi, size=(edges_per_robot, n)).astype(np.float32)
        u = rng.normal(size=(edges_per_robot, n))
        u = (u / np.linalg.norm(u, axis=1, keepdims=True)).astype(np.float32)
        K = np.rint(np.exp(rng.uniform(np.log(8), np.log(256), edges_per_robot))).astype(np.int32)
        batches.append(EdgeBatch(rb, q1, u, K))
    return EdgeWorkload(sc, batches)


def _capsule_overlaps_box(P: np.ndarray, radius, aabb: np.ndarray, margin: float = 0.0) -> bool:
    """Conservative test: some capsule of the skeleton P ((L + 1) x 3,
    radii per link) has its bounds grown by radius + margin overlapping the
    box's bounds."""
    for i in range(len(P) - 1):
        g = radius[i] + margin
        lo = np.minimum(P[i], P[i + 1]) - g
        hi = np.maximum(P[i], P[i + 1]) + g
        if np.all(aabb[:3] <= hi) and np.all(aabb[3:] >= lo):
            return True
    return False


def workload3_planner(edges_per_robot: int = 4096, dofs=(10, 14, 18, 22), spread: float = 0.35,
                      oracle=None, judge=None) -> EdgeWorkload:
    """Edges as a tree-based planner checks them: RRTConnect::extend checks
    only edges that leave a tree node, and tree nodes are collision-free
    (RRTConnect.cpp:92-105 -> RealVectorSpace.cpp:261-275; the scene and
    start rule of CommonFunctions.h:77-113, which re-draws every random
    obstacle that makes the start invalid).  As specified, workload 3 has no
    such state: one of its 64 boxes overlaps the chains' fixed base link, so
    every configuration collides at state 0.

    Robots: workload 3's four chains (10/14/18/22 DOF, r = 0.04,
    self-collision on).  Scene: 64 boxes drawn as in workload 3 (sides
    U[0.1, 0.5], centres uniform in a 1.5 m ball), each re-drawn while its
    bounds come within 5 cm of the bounds of a capsule of any chain's home
    pose (all joints 0: the chain straight up from the origin).  Starts: joint 1 uniform in [-pi, pi],
    joints 2..n uniform in [-spread, spread] rad around the home pose,
    re-drawn until the start is collision-free against the boxes and itself
    (the oracle's own edge predicate, a one-state edge).  Edges: as in
    workload 3 (a uniform direction in joint space, K = round(exp U[ln 8,
    ln 256]) states 0.05 rad apart).  The start states are the only change
    to the distribution; uniform starts in [-pi, pi]^n are collision-free
    with probability below 1e-6 on the 22-DOF chain (6e-4 for its
    self-collision alone).

    `oracle` is the module that judges the starts (pyoracle): the generator
    is test and benchmark infrastructure.  Or `judge(robot, scene, q)`, the
    collision-free mask of the one-state edges q by another predicate with
    the same verdicts (the engine's reference-arithmetic edge search, for
    programs that do not load the oracle): the same starts."""
    if oracle is None and judge is None:
        raise ValueError("workload3_planner needs the oracle module or a judge of start states")
    robots = [b.robot for b in workload3(16, dofs).batches]
    homes = [R.skeleton_np(rb, np.zeros(rb.num_dofs, np.float32))[0] for rb in robots]
    rng = np.random.default_rng(1013)
    kept: List[Scene] = []
    while len(kept) < 64:
        b = _boxes(rng, 1, 0.1, 0.5, 1.5)
        if any(_capsule_overlaps_box(P, rb.radius, b.aabb[0], 0.05) for P, rb in zip(homes, robots)):
            continue
        kept.append(b)
    sc = kept[0]
    for b in kept[1:]:
        sc = sc.concat(b)
    if judge is None:
        osc = oracle.scene_of(sc)

        def judge(rb, scene, q):
            return oracle.edge_first_collision(rb, osc, q, np.zeros_like(q), np.ones(len(q), np.int32), 0.05) < 0
    batches = []
    for rb in robots:
        n = rb.num_dofs
        starts = np.zeros((0, n), np.float32)
        while len(starts) < edges_per_robot:
            m = 4 * edges_per_robot
            q = np.empty((m, n), np.float32)
            q[:, 0] = rng.uniform(-np.pi, np.pi, m)
            q[:, 1:] = rng.uniform(-spread, spread, (m, n - 1))
            starts = np.concatenate([starts, q[judge(rb, sc, q)]])
        q1 = np.ascontiguousarray(starts[:edges_per_robot])
        u = rng.normal(size=(edges_per_robot, n))
        u = (u / np.linalg.norm(u, axis=1, keepdims=True)).astype(np.float32)
        K = np.rint(np.exp(rng.uniform(np.log(8), np.log(256), edges_per_robot))).astype(np.int32)
        batches.append(EdgeBatch(rb, q1, u, K))
    return EdgeWorkload(sc, batches)


# ---------------------------------------------------------------- workload 4
@dataclass
class ServingWorkload:
    """Request stream for the serving loop.  Robots 0 = xArm6 (r = 0.05,
    ground from link 1), 1 = the repository's spatial 10-DOF chain
    (r = 0.025, ground from link 2); the scene is workload 2's (ground + 16
    boxes); each of the 16 boxes moves at a constant velocity s*u,
    s ~ U[0, 0.5] m/s, u a uniform direction.  Requests arrive as a Poisson
    process at `rate` per second; sizes are floor(exp U[0, ln 65536]); robot
    and kind (0 verdict, 1 underestimation) are 50/50.  Each request's
    configurations are a slice [offset, offset + n) of its robot's pool."""
    robots: List[R.Robot]
    scene: Scene
    vel: np.ndarray          # (M, 3) float32, 0 for the ground
    pools: List[np.ndarray]  # per robot (P, dofs) float32
    arrival: np.ndarray      # (K,) seconds from the start
    robot: np.ndarray        # (K,) int32
    kind: np.ndarray         # (K,) int32
    size: np.ndarray         # (K,) int64
    offset: np.ndarray       # (K,) int64
    tick_seconds: float = 0.01


def workload4(duration_s: float = 5.0, rate: float = 20000.0, pool: int = 1 << 21) -> ServingWorkload:
    rng = np.random.default_rng(1004)
    sc = scene2(np.random.default_rng(1002))
    robots = [R.xarm6(radius=0.05, ground_included=1), R.spatial(10, radius=0.025, ground_included=2)]
    M = sc.num_obj
    speed = rng.uniform(0.0, 0.5, M)
    u = rng.normal(size=(M, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    vel = (speed[:, None] * u).astype(np.float32)
    vel[sc.is_ground == 1] = 0.0
    pools = [uniform_configs(rng, pool, rb.limits) for rb in robots]
    gaps = rng.exponential(1.0 / rate, size=int(duration_s * rate * 1.2) + 16)
    arrival = np.cumsum(gaps)
    arrival = arrival[arrival < duration_s]
    K = len(arrival)
    size = np.floor(np.exp(rng.uniform(0.0, np.log(65536.0), K))).astype(np.int64)
    size = np.clip(size, 1, 65536)
    robot = rng.integers(0, 2, K).astype(np.int32)
    kind = rng.integers(0, 2, K).astype(np.int32)
    offset = np.floor(rng.random(K) * (pool - size + 1)).astype(np.int64)
    return ServingWorkload(robots, sc, vel, pools, arrival, robot, kind, size, offset)


def scene_at_tick(scene: Scene, vel: np.ndarray, tick: int, tick_seconds: float) -> Scene:
    """Box positions the serving loop uses at a tick:
    pos = pos0 + vel * (tick * tick_seconds), in float32 (no contraction)."""
    t = np.float32(tick) * np.float32(tick_seconds)
    pos = (scene.pos + np.asarray(vel, np.float32) * t).astype(np.float32)
    return Scene(scene.dim, pos, scene.quat, scene.is_ground, scene.min_dist_tol)


# ---------------------------------------------------------------- workload 5
def workload5(n: int = 16777216, rotated_semantics: bool = True) -> ConfigBatch:
    rng = np.random.default_rng(1005)
    sc = _boxes(rng, 512, 0.02, 0.2, 2.0, rotated=True)
    lengths = rng.uniform(0.1, 0.3, 22)
    rb = R.chain(lengths, radius=0.05, name="chain22", self_collision=True)
    q = rng.uniform(-np.pi, np.pi, size=(n, 22)).astype(np.float32)
    return ConfigBatch("chain22_512obb", rb, sc, q)


# ---------------------------------------------------------------- workload 6
@dataclass
class BurWorkload:
    """Generalized spines from one root towards many targets (a generalized
    bur, RGBTConnect / DRGBT), with the planner's spine settings."""
    name: str
    robot: R.Robot
    scene: Scene
    root: np.ndarray        # (n,) float32
    q_e: np.ndarray         # (N, n) float32
    num_layers: int
    num_iter: int
    d_crit: float
    use_real_dist: bool


def workload6(n: int = 65536, use_real_dist: bool = False) -> BurWorkload:
    """xArm6 in the ground + 16-box scene of workload 2, root at q = 0 (clear
    of every box by construction, 0.062 m above the ground), targets at
    DELTA = pi rad in uniform random directions; NUM_LAYERS 5,
    NUM_ITER_SPINE 1, D_CRIT 0.05 (data/configurations/
    configuration_rgbtconnect.yaml, configuration_rbtconnect.yaml)."""
    sc = scene2(np.random.default_rng(1002))
    rb = R.xarm6(radius=0.05, ground_included=1)
    rng = np.random.default_rng(1006)
    root = np.zeros(6, np.float32)
    u = rng.normal(size=(n, 6))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    q_e = (root + np.pi * u).astype(np.float32)
    return BurWorkload("xarm6_gbur_16aabb_ground", rb, sc, root, q_e, 5, 1, 0.05, use_real_dist)
