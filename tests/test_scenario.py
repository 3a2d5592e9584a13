"""Scenario files read by rpmplv2_amd.scenario (CPU).

Fixtures: tests/golden/scenarios/*.yaml hold the data of RPMPLv2 scenario
files (comments removed) and the first runs of a random_scenarios.yaml;
tests/golden/reference_scenarios.json holds the same scenarios as parsed
independently when the fixtures were made.
"""
import json
import os

import numpy as np
import pytest

from rpmplv2_amd import robots as R, scenario as S, workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
SC = os.path.join(HERE, "golden", "scenarios")


def _json():
    with open(os.path.join(HERE, "golden", "reference_scenarios.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("robot,name", [("planar_2dof", "scenario_test"), ("planar_2dof", "scenario1"),
                                        ("spatial_10dof", "scenario1"), ("xarm6", "scenario_test"),
                                        ("xarm6", "scenario1"), ("xarm6", "scenario2")])
def test_scenarios_match_the_reference_data(oracle, robot, name):
    """Robot, boxes (ground rule, (x, y, z, w) rotations, tolerances), start
    and goal as in the independently parsed record; start and goal are
    valid (the planners refuse invalid ones, RRTConnect.cpp:13-16)."""
    sc = S.load_scenario(os.path.join(SC, f"{robot}_{name}.yaml"))
    rec = next(f for f in _json()["fixed"] if f["robot"] == robot and f["scenario"] == name)
    assert sc.robot.num_dofs == rec["num_dofs"]
    assert sc.robot.ground_included == rec["ground_included"]
    boxes = [b for b in rec["boxes"] if not (b["label"] == "ground" and rec["ground_included"] == 0)]
    assert sc.scene.num_obj == len(boxes)
    for j, b in enumerate(boxes):
        assert np.allclose(sc.scene.dim[j], b["dim"]) and np.allclose(sc.scene.pos[j], b["pos"])
        q = [1, 0, 0, 0] if b["rot"] is None else [b["rot"][3], b["rot"][0], b["rot"][1], b["rot"][2]]
        assert np.allclose(sc.scene.quat[j], q)
        assert sc.scene.min_dist_tol[j] == np.float32(b["min_dist_tol"])
        assert sc.scene.is_ground[j] == (b["label"] == "ground")
    assert np.allclose(sc.q_start, rec["q_start"]) and np.allclose(sc.q_goal, rec["q_goal"])
    if rec["capsules_radius"]:
        r = rec["capsules_radius"]
        assert np.allclose(sc.robot.radius, r if len(r) > 1 else r * sc.robot.num_dofs)
    scene = oracle.scene_of(sc.scene)
    assert oracle.is_valid(sc.robot, scene, np.stack([sc.q_start, sc.q_goal])).all()


def test_planar_10dof_scenario(oracle):
    sc = S.load_scenario(os.path.join(SC, "planar_10dof_scenario_test.yaml"))
    assert sc.robot.name == "planar_10dof" and sc.robot.num_dofs == 10
    assert sc.space == "RealVectorSpaceFCL" and sc.scene.num_obj == 2
    P = oracle.skeleton(sc.robot, np.zeros((1, 10), np.float32), 64)[0]
    assert np.allclose(P[:, 0], 0.2 * np.arange(11)) and np.allclose(P[:, 1:], 0)   # 0.2 m links along x
    assert np.allclose(sc.robot.radius, 0.025)                                       # URDF box width / 2


def test_random_scenarios_through_the_loader(oracle):
    """Runs read from random_scenarios.yaml into the scenario's scene: each
    start is >= 0.3 m from its run's boxes (6 * D_CRIT,
    apps/CommonFunctions.h:159-160; printed with 6 digits: 3e-4 m slack)."""
    sc = S.load_scenario(os.path.join(SC, "spatial_10dof_scenario_random_obstacles.yaml"))
    runs = S.load_random_scenarios(os.path.join(SC, "spatial_10dof_random_scenarios.yaml"))
    assert len(runs) == 5 and all(len(r.obstacle_pos) == len(r.obstacle_vel) > 0 for r in runs)
    rec = _json()["random"]["spatial_10dof"]["runs"]
    for run in runs:
        same = next(x for x in rec if x["scenario"] == run.scenario and x["run"] == run.run)
        assert np.allclose(run.start, same["start"]) and np.allclose(run.obstacle_pos, same["objects"])
        moving = W.Scene(np.tile(sc.random_obstacles["dim"], (len(run.obstacle_pos), 1)), run.obstacle_pos,
                         np.tile([1.0, 0, 0, 0], (len(run.obstacle_pos), 1)), np.zeros(len(run.obstacle_pos)),
                         np.full(len(run.obstacle_pos), W.MAX_DISTANCE))
        d, _, _, _ = oracle.distance(sc.robot, oracle.scene_of(moving), run.start[None], 32)
        assert d[0] >= 0.3 - 3e-4
        full = S.run_scene(sc, run)
        assert full.num_obj == sc.scene.num_obj + len(run.obstacle_pos)
        assert (full.dynamic[sc.scene.num_obj:] == 1).all()


TEXT = """environment:
\t- box:
\t    label: "ground"
\t    dim: [1.5, 1.5, 0.1]
\t    pos: [0, 0, -0.05]
\t- box:
\t    dim: [0.2, 0.4, 0.1]
\t    pos: [0.5, 0, 0.3]
\t    rot: [0, 0, 0.3826834, 0.9238795]
\t    min_dist_tol: 0.1
robot:
  type: "spatial_10dof"
  num_DOFs: 10
  capsules_radius: [0.03]
  max_vel: [1.5]
  ground_included: GI
  q_start: [0, 0, 0, 0, 0, 0, 0, 0, 0, 0]
  q_goal: [1, 0, 0, 0, 0, 0, 0, 0, 0, 0]
testing:
  motion_type: "light_directions"
"""


def test_scenario_rules():
    sc = S.parse_scenario(TEXT.replace("GI", "2"))
    assert sc.scene.num_obj == 2 and sc.scene.is_ground.tolist() == [1, 0]
    assert np.allclose(sc.scene.quat[1], [0.9238795, 0, 0, 0.3826834])      # rot is (x, y, z, w)
    assert sc.scene.min_dist_tol.tolist() == [W.MAX_DISTANCE, np.float32(0.1)]
    assert np.allclose(sc.robot.radius, 0.03) and np.allclose(sc.max_vel, 1.5)   # one value for every joint
    assert sc.testing["motion_type"] == "light_directions"
    sc0 = S.parse_scenario(TEXT.replace("GI", "0"))
    assert sc0.scene.num_obj == 1 and sc0.scene.is_ground.tolist() == [0]   # ground dropped
    with pytest.raises(S.ScenarioError):
        S.parse_scenario(TEXT.replace("GI", "2").replace("q_goal: [1, 0, 0, 0, 0, 0, 0, 0, 0, 0]", "q_goal: [1, 0]"))
    with pytest.raises(S.ScenarioError):
        S.parse_scenario(TEXT.replace("GI", "2").replace("[0.03]", "[0.03, 0.04]"))
    with pytest.raises(S.ScenarioError):
        S.parse_scenario(TEXT.replace("GI", "2").replace("spatial_10dof", "unknown_robot"))
    xarm = TEXT.replace("spatial_10dof", "xarm6").replace("num_DOFs: 10", "num_DOFs: 6").replace(
        "  ground_included: GI\n", "").replace("0, 0, 0, 0, 0, 0, 0, 0, 0, 0", "0, 0, 0, 0, 0, 0").replace(
        "1, 0, 0, 0, 0, 0, 0, 0, 0, 0", "1, 0, 0, 0, 0, 0")
    with pytest.raises(S.ScenarioError):
        S.parse_scenario(xarm)          # xarm6 must say whether the ground is included
