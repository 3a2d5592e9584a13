"""Robot descriptions for the capsule collision engine.

A robot is a serial chain of URDF-style joints (origin xyz/rpy, axis, type)
from the base link to the tool, a skeleton recipe that turns the chain's
frames into capsule end points, one capsule radius per link, the number of
leading links that ignore the "ground" obstacle, and an optional list of
link pairs checked for self-collision.

The kinematic facts below restate RPMPLv2's robot files:
  xArm6            data/xarm6/xarm6.urdf:69-222 (chain link_base -> link_eef,
                   xArm6.cpp:21); skeleton recipe xArm6.cpp:209-235
  planar 2-DOF     data/planar_2dof/planar_2dof.urdf:50-72
  spatial n-DOF    data/spatial_{10,14,18,22}dof/spatial_*dof.urdf (joint 1
                   about z at the origin, joint 2 at z=0.1 about x, the rest
                   and the tool at z=L about x); skeleton = segment origins,
                   Planar2DOF.cpp:117-130
Capsule radii default to the scenario values the reference ships
(Scenario.cpp:36-51 broadcasts a single value to every link).
"""
from __future__ import annotations

import math
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

JOINT_FIXED = 0
JOINT_REVOLUTE = 1
SKEL_CHAIN = 0
SKEL_XARM6 = 1


@dataclass
class Joint:
    type: int
    xyz: Tuple[float, float, float]
    rpy: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    axis: Tuple[float, float, float] = (0.0, 0.0, 1.0)
    lower: float = -math.pi
    upper: float = math.pi


@dataclass
class Robot:
    name: str
    skeleton_kind: int
    joints: List[Joint]
    num_links: int
    radius: List[float]
    ground_included: int = 0
    gripper_length: float = 0.0
    self_pairs: List[Tuple[int, int]] = field(default_factory=list)

    @property
    def num_dofs(self) -> int:
        return sum(1 for j in self.joints if j.type == JOINT_REVOLUTE)

    @property
    def limits(self) -> np.ndarray:
        return np.array([(j.lower, j.upper) for j in self.joints if j.type == JOINT_REVOLUTE],
                        dtype=np.float32)

    def with_radius(self, r) -> "Robot":
        rr = [float(r)] * self.num_links if np.isscalar(r) else [float(x) for x in r]
        if len(rr) != self.num_links:
            raise ValueError("number of capsule radii does not match the number of links")
        return Robot(self.name, self.skeleton_kind, list(self.joints), self.num_links, rr,
                     self.ground_included, self.gripper_length, list(self.self_pairs))


def xarm6(radius: float | Sequence[float] = 0.05, ground_included: int = 1,
          gripper_length: float = 0.0) -> Robot:
    """UFACTORY xArm6 (rpy uses the literal 1.5708 of the robot file)."""
    h = 1.5708
    lim = 3.14159265359
    joints = [
        Joint(JOINT_REVOLUTE, (0, 0, 0.267), (0, 0, 0), (0, 0, 1), -lim, lim),
        Joint(JOINT_REVOLUTE, (0, 0, 0), (-h, 0, 0), (0, 0, 1), -2.059, 2.0944),
        Joint(JOINT_REVOLUTE, (0.0535, -0.2845, 0), (0, 0, 0), (0, 0, 1), -3.927, 0.19198),
        Joint(JOINT_REVOLUTE, (0.0775, 0.3425, 0), (-h, 0, 0), (0, 0, 1), -lim, lim),
        Joint(JOINT_REVOLUTE, (0, 0, 0), (h, 0, 0), (0, 0, 1), -1.69297, lim),
        Joint(JOINT_REVOLUTE, (0.076, 0.097, 0), (-h, 0, 0), (0, 0, 1), -lim, lim),
        Joint(JOINT_FIXED, (0, 0, 0)),
    ]
    r = Robot("xarm6", SKEL_XARM6, joints, 6, [0.0] * 6, ground_included, gripper_length,
              [(0, 5), (1, 5), (0, 4), (1, 4)])
    return r.with_radius(radius)


def planar_2dof(radius: float | Sequence[float] = 0.025) -> Robot:
    joints = [
        Joint(JOINT_REVOLUTE, (0, 0, 0), axis=(0, 0, 1), lower=-3.14159, upper=3.14159),
        Joint(JOINT_REVOLUTE, (1, 0, 0), axis=(0, 0, 1), lower=-3.14159, upper=3.14159),
        Joint(JOINT_FIXED, (1, 0, 0)),
    ]
    return Robot("planar_2dof", SKEL_CHAIN, joints, 2, [0.0, 0.0]).with_radius(radius)


def planar_10dof(radius: float | Sequence[float] = 0.025) -> Robot:
    """The repository's planar_10dof robot: ten revolute joints about z,
    0.2 m apart along x, tool 0.2 m past the last joint."""
    joints = [Joint(JOINT_REVOLUTE, (0, 0, 0), axis=(0, 0, 1), lower=-3.14159, upper=3.14159)]
    for _ in range(9):
        joints.append(Joint(JOINT_REVOLUTE, (0.2, 0, 0), axis=(0, 0, 1), lower=-3.14159, upper=3.14159))
    joints.append(Joint(JOINT_FIXED, (0.2, 0, 0)))
    return Robot("planar_10dof", SKEL_CHAIN, joints, 10, [0.0] * 10).with_radius(radius)


_SPATIAL_L = {10: 0.1, 14: 0.0692, 18: 0.05294, 22: 0.042857}


def chain(lengths: Sequence[float], radius: float | Sequence[float] = 0.025,
          ground_included: int = 0, name: str = "chain", self_collision: bool = False) -> Robot:
    """Spatial chain: joint 1 about z at the origin, joints 2..n about x,
    joint k+1 (and the tool) offset by lengths[k-1] along z."""
    n = len(lengths)
    joints = [Joint(JOINT_REVOLUTE, (0, 0, 0), axis=(0, 0, 1), lower=-3.14159, upper=3.14159)]
    for k in range(1, n):
        joints.append(Joint(JOINT_REVOLUTE, (0, 0, float(lengths[k - 1])), axis=(1, 0, 0),
                            lower=-3.14159, upper=3.14159))
    joints.append(Joint(JOINT_FIXED, (0, 0, float(lengths[n - 1]))))
    pairs = [(i, j) for i in range(n) for j in range(i + 2, n)] if self_collision else []
    return Robot(name, SKEL_CHAIN, joints, n, [0.0] * n, ground_included, 0.0, pairs).with_radius(radius)


def spatial(n: int, radius: float | Sequence[float] = 0.025, ground_included: int = 2,
            self_collision: bool = False) -> Robot:
    """The repository's spatial_{10,14,18,22}dof robots."""
    L = _SPATIAL_L[n]
    return chain([0.1] + [L] * (n - 1), radius, ground_included, f"spatial_{n}dof", self_collision)


# ---------------------------------------------------------------- URDF subset
def load_urdf(path: str, base: str, tip: str, skeleton_kind: int = SKEL_CHAIN,
              radius: float | Sequence[float] = 0.025, num_links: Optional[int] = None,
              ground_included: int = 0, gripper_length: float = 0.0) -> Robot:
    """Read the joint chain base->tip from a URDF file (joints only)."""
    root = ET.parse(path).getroot()
    by_child = {}
    for j in root.findall("joint"):
        by_child[j.find("child").get("link")] = j
    chain_j = []
    link = tip
    while link != base:
        j = by_child.get(link)
        if j is None:
            raise ValueError(f"no joint leads to link {link!r}")
        chain_j.append(j)
        link = j.find("parent").get("link")
    chain_j.reverse()
    joints = []
    for j in chain_j:
        o = j.find("origin")
        xyz = tuple(float(v) for v in (o.get("xyz", "0 0 0") if o is not None else "0 0 0").split())
        rpy = tuple(float(v) for v in (o.get("rpy", "0 0 0") if o is not None else "0 0 0").split())
        a = j.find("axis")
        axis = tuple(float(v) for v in (a.get("xyz") if a is not None else "1 0 0").split())
        lim = j.find("limit")
        lo = float(lim.get("lower", -math.pi)) if lim is not None else -math.pi
        hi = float(lim.get("upper", math.pi)) if lim is not None else math.pi
        t = JOINT_REVOLUTE if j.get("type") in ("revolute", "continuous") else JOINT_FIXED
        joints.append(Joint(t, xyz, rpy, axis, lo, hi))
    ndof = sum(1 for j in joints if j.type == JOINT_REVOLUTE)
    nl = num_links if num_links is not None else ndof
    return Robot(root.get("name", "robot"), skeleton_kind, joints, nl, [0.0] * nl,
                 ground_included, gripper_length).with_radius(radius)


# ---------------------------------------------------------------- host FK
def _rpy_matrix(rpy) -> np.ndarray:
    r, p, y = rpy
    cr, sr, cp, sp, cy, sy = math.cos(r), math.sin(r), math.cos(p), math.sin(p), math.cos(y), math.sin(y)
    return np.array([[cy * cp, cy * sp * sr - sy * cr, cy * sp * cr + sy * sr],
                     [sy * cp, sy * sp * sr + cy * cr, sy * sp * cr - cy * sr],
                     [-sp, cp * sr, cp * cr]])


def skeleton_np(rb: Robot, q) -> np.ndarray:
    """Capsule end points in double for host-side scene set-up (obstacle
    rejection, workspace checks).  Not used on the query path."""
    q = np.atleast_2d(np.asarray(q, np.float64))
    n = q.shape[0]
    T_R = np.broadcast_to(np.eye(3), (n, 3, 3)).copy()
    T_p = np.zeros((n, 3))
    frames = []
    qi = 0
    for j in rb.joints:
        R0 = _rpy_matrix(j.rpy)
        if j.type == JOINT_REVOLUTE:
            a = R0 @ np.asarray(j.axis, np.float64)
            a = a / np.linalg.norm(a)
            th = q[:, qi]; qi += 1
            c, s = np.cos(th), np.sin(th)
            K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
            Rq = (np.eye(3)[None] + s[:, None, None] * K[None] + (1 - c)[:, None, None] * (K @ K)[None])
            S_R = Rq @ R0
        else:
            S_R = np.broadcast_to(R0, (n, 3, 3))
        T_p = T_p + np.einsum("nij,j->ni", T_R, np.asarray(j.xyz, np.float64))
        T_R = T_R @ S_R
        frames.append((T_R.copy(), T_p.copy()))
    L = rb.num_links
    P = np.zeros((n, L + 1, 3))
    if rb.skeleton_kind == SKEL_CHAIN:
        for k in range(L + 1):
            P[:, k] = frames[k][1]
    else:
        f = frames
        g = rb.gripper_length
        p5 = f[5][1] + g * f[5][0][:, :, 2]
        P[:, 1] = f[1][1]
        P[:, 2] = f[2][1]
        P[:, 3] = f[3][1] - 0.25 * f[3][0][:, :, 2]
        P[:, 4] = f[4][1]
        P[:, 5] = f[4][1] + 0.076 * f[4][0][:, :, 0]
        P[:, 6] = p5 - 0.3 * g * f[5][0][:, :, 2]
    return P
