"""Scenario files of RPMPLv2 read into this engine's robot and scene types.

Two formats:
* scenario YAML (scenario::Scenario, src/scenario/Scenario.cpp:3-128, and
  env::Environment, src/environments/Environment.cpp:13-109): the robot node
  (type, num_DOFs, capsules_radius, q_start, q_goal, ground_included,
  gripper_length, self_collision_checking, WS_center, WS_radius, max_vel /
  max_acc / max_jerk), the environment's boxes, random_obstacles and testing;
* random_scenarios.yaml (apps/generate_random_scenarios.cpp:43-96): per
  scenario and run a start, a goal and the random dynamic obstacles' positions
  and velocities.

The rules follow the reference loaders: a box labelled "ground" is dropped
when ground_included is 0 (Environment.cpp:32-33); rot is read as
(x, y, z, w) (:44-45); min_dist_tol defaults to MAX_DISTANCE (Box.cpp:13); a
single capsule radius or joint limit value applies to every joint
(Scenario.cpp:36-51, 53-97); start and goal must have num_DOFs entries
(:114-121).  yaml-cpp accepts tabs as separators, so tabs are expanded before
parsing.  The robot's kinematics come from this package's tables
(rpmplv2_amd.robots), or from the URDF when a root path is given and the
type has no table.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import robots as R
from .workloads import MAX_DISTANCE, Scene


class ScenarioError(ValueError):
    pass


@dataclass
class Scenario:
    robot: R.Robot
    scene: Scene
    q_start: np.ndarray
    q_goal: np.ndarray
    space: str = "RealVectorSpace"
    self_collision_checking: bool = False
    ws_center: Optional[np.ndarray] = None
    ws_radius: Optional[float] = None
    max_vel: Optional[np.ndarray] = None
    max_acc: Optional[np.ndarray] = None
    max_jerk: Optional[np.ndarray] = None
    random_obstacles: Dict = field(default_factory=dict)
    testing: Dict = field(default_factory=dict)


@dataclass
class RandomRun:
    scenario: str
    run: str
    start: np.ndarray
    goal: np.ndarray
    obstacle_pos: np.ndarray   # (K, 3)
    obstacle_vel: np.ndarray   # (K, 3)


def _yaml(text: str):
    import yaml
    return yaml.safe_load(text.replace("\t", "    "))


def _per_dof(node, n: int, what: str) -> Optional[np.ndarray]:
    if node is None:
        return None
    vals = [float(v) for v in node]
    if len(vals) == 1:
        return np.full(n, vals[0], np.float32)
    if len(vals) != n:
        raise ScenarioError(f"the size of '{what}' is not correct")
    return np.asarray(vals, np.float32)


def _robot(node, root_path: Optional[str]) -> R.Robot:
    kind = str(node["type"])
    n = int(node["num_DOFs"])
    gi = node.get("ground_included")
    radius = _per_dof(node.get("capsules_radius"), n, "capsules_radius")
    if kind == "xarm6":
        if gi is None:
            raise ScenarioError("it is not defined whether the ground is included")
        if radius is None:
            raise ScenarioError("xarm6 needs capsules_radius")
        return R.xarm6(radius=list(radius), ground_included=int(gi),
                       gripper_length=float(node.get("gripper_length", 0.0)))
    r = 0.025 if radius is None else list(radius)
    if kind == "planar_2dof":
        rb = R.planar_2dof(radius=r)
    elif kind == "planar_10dof":
        rb = R.planar_10dof(radius=r)
    elif kind.startswith("spatial_") and kind.endswith("dof"):
        rb = R.spatial(int(kind[len("spatial_"):-3]), radius=r, ground_included=int(gi or 0))
    elif root_path and node.get("urdf"):
        rb = R.load_urdf(os.path.join(root_path, str(node["urdf"]).lstrip("/")), "base_link", "tool", radius=r,
                         ground_included=int(gi or 0))
    else:
        raise ScenarioError(f"robot type {kind!r} is not known")
    if rb.num_dofs != n:
        raise ScenarioError(f"num_DOFs {n} does not match the {kind} robot")
    return rb


def _scene(env_node, ground_included: int) -> Scene:
    dim, pos, quat, ground, tol, dyn = [], [], [], [], [], []
    for obj in env_node or []:
        if "box" not in obj:
            if "sphere" in obj:
                continue           # the reference reads no sphere either (Environment.cpp:53-56)
            raise ScenarioError("object type is wrong")
        b = obj["box"]
        label = str(b.get("label", ""))
        if label == "ground" and ground_included == 0:
            continue
        dim.append([float(v) for v in b["dim"]])
        pos.append([float(v) for v in b["pos"]])
        r = b.get("rot")
        quat.append([float(r[3]), float(r[0]), float(r[1]), float(r[2])] if r is not None else [1.0, 0.0, 0.0, 0.0])
        ground.append(1 if label == "ground" else 0)
        dyn.append(1 if label == "dynamic_obstacle" else 0)
        tol.append(float(b["min_dist_tol"]) if "min_dist_tol" in b else MAX_DISTANCE)
    m = len(dim)
    return Scene(np.reshape(dim, (m, 3)), np.reshape(pos, (m, 3)), np.reshape(quat, (m, 4)), ground, tol,
                 dynamic=np.asarray(dyn, np.uint8))


def parse_scenario(text: str, root_path: Optional[str] = None) -> Scenario:
    """A scenario YAML document (the text of a data/*/scenario*/*.yaml file)."""
    node = _yaml(text)
    rn = node["robot"]
    n = int(rn["num_DOFs"])
    rb = _robot(rn, root_path)
    q_start, q_goal = rn.get("q_start"), rn.get("q_goal")
    if q_start is None or q_goal is None or len(q_start) != n or len(q_goal) != n:
        raise ScenarioError("start or goal size is not correct")
    default_self = rn["type"] == "xarm6"
    return Scenario(
        robot=rb, scene=_scene(node.get("environment"), rb.ground_included),
        q_start=np.asarray(q_start, np.float32), q_goal=np.asarray(q_goal, np.float32),
        space=str(rn.get("space", "RealVectorSpace")),
        self_collision_checking=bool(rn.get("self_collision_checking", default_self)),
        ws_center=np.asarray(rn["WS_center"], np.float32) if "WS_center" in rn else None,
        ws_radius=float(rn["WS_radius"]) if "WS_radius" in rn else None,
        max_vel=_per_dof(rn.get("max_vel"), n, "max_vel"),
        max_acc=_per_dof(rn.get("max_acc"), n, "max_acc"),
        max_jerk=_per_dof(rn.get("max_jerk"), n, "max_jerk"),
        random_obstacles=dict(node.get("random_obstacles") or {}),
        testing=dict(node.get("testing") or {}))


def load_scenario(path: str, root_path: Optional[str] = None) -> Scenario:
    with open(path) as f:
        return parse_scenario(f.read(), root_path)


def parse_random_scenarios(text: str) -> List[RandomRun]:
    """random_scenarios.yaml: scenario_<k> -> run_<i> -> start, goal,
    object_<j> {pos, vel}, in file order."""
    node = _yaml(text) or {}
    out = []
    for sname, runs in node.items():
        for rname, run in (runs or {}).items():
            objs = [run[k] for k in run if str(k).startswith("object_")]
            out.append(RandomRun(str(sname), str(rname), np.asarray(run["start"], np.float32),
                                 np.asarray(run["goal"], np.float32),
                                 np.asarray([o["pos"] for o in objs], np.float32).reshape(-1, 3),
                                 np.asarray([o.get("vel", [0, 0, 0]) for o in objs], np.float32).reshape(-1, 3)))
    return out


def load_random_scenarios(path: str) -> List[RandomRun]:
    with open(path) as f:
        return parse_random_scenarios(f.read())


def run_scene(sc: Scenario, run: RandomRun) -> Scene:
    """The scenario's boxes plus a run's random obstacles: boxes of
    random_obstacles.dim at the stored positions, labelled dynamic_obstacle
    (apps/CommonFunctions.h:77-177)."""
    dim = np.asarray(sc.random_obstacles.get("dim", [0.01, 0.01, 0.01]), np.float32)
    k = len(run.obstacle_pos)
    s = sc.scene
    dyn0 = s.dynamic if s.dynamic is not None else np.zeros(s.num_obj, np.uint8)
    return Scene(np.concatenate([s.dim, np.tile(dim, (k, 1))]), np.concatenate([s.pos, run.obstacle_pos]),
                 np.concatenate([s.quat, np.tile([1.0, 0, 0, 0], (k, 1))]),
                 np.concatenate([s.is_ground, np.zeros(k, np.uint8)]),
                 np.concatenate([s.min_dist_tol, np.full(k, MAX_DISTANCE, np.float32)]),
                 dynamic=np.concatenate([dyn0, np.ones(k, np.uint8)]))
