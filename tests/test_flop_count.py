"""The operation-counting restatement (oracle/oracle_flops.h) behind the VALU roofline figures (profiles/flops_*.json):
it must run the FP64 restatement's arithmetic unchanged, and its counts must behave as operation counts do."""

from __future__ import annotations

import json
import os

import numpy as np
import pytest

from genesis_amd.model.compiler import SolverOptions, compile_model
from genesis_amd.model.primitives import box_entity, plane_entity
from genesis_amd.runtime import STAGE_ID
from genesis_amd.scenes import franka_initial_qpos, franka_model, franka_targets
from tests.oracle_lib import Oracle, flops

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _franka(precision, B):
    m = franka_model()
    o = Oracle(m, B, precision=precision)
    o.set("qpos", franka_initial_qpos(m, B, 0))
    o.set("ctrl_mode", np.zeros((B, m.n_dofs)))
    o.forward_kinematics()
    o.set("ctrl_pos", franka_targets(m, B, 0))
    return o


def test_counting_type_keeps_the_fp64_arithmetic():
    a, b = _franka(64, 4), _franka(0, 4)
    a.step(30, 1)
    f0 = flops()[0]
    b.step(30, 1)
    assert flops()[0] > f0
    for name in ("qpos", "vel", "acc", "contact_pos"):
        assert np.array_equal(a.get(name), b.get(name)), name


def test_counts_add_over_environments():
    one, four = _franka(0, 1), _franka(0, 4)
    for o in (one, four):  # the same state in every environment
        o.set("qpos", np.repeat(franka_initial_qpos(o.model, 1, 0), o.n_envs, 0))
        o.set("ctrl_pos", np.repeat(franka_targets(o.model, 1, 0), o.n_envs, 0))
        o.forward_kinematics()
    f0 = flops()[0]
    one.step(5, 1)
    f1 = flops()[0]
    four.step(5, 1)
    f2 = flops()[0]
    assert f2 - f1 == 4 * (f1 - f0)


def _boxes_in_the_air(n):
    ents = [box_entity([0.1, 0.2, 0.3], pos=(3.0 * i, 0, 2.0), name=f"b{i}") for i in range(n)]
    m = compile_model(ents, SolverOptions())
    return Oracle(m, 1, precision=0)


@pytest.mark.parametrize("stage", ["crba", "factor", "rne", "integrate"])
def test_dynamics_counts_grow_linearly_with_independent_bodies(stage):
    """N free boxes far apart: the mass matrix is block diagonal, so the dynamics stages do the same work per body
    (the second difference of the count over N = 1, 2, 3 is zero; a dense D x D pass would make it positive)."""
    counts = []
    for n in (1, 2, 3):
        o = _boxes_in_the_air(n)
        o.forward_kinematics()
        f0 = flops()[0]
        o.stage(STAGE_ID[stage])
        counts.append(flops()[0] - f0)
    assert counts[0] > 0
    assert counts[2] - 2 * counts[1] + counts[0] == 0, counts


@pytest.mark.parametrize("name", ["flops_franka4096.json", "flops_boxpile1024.json", "flops_quadruped8192.json"])
def test_committed_counts_are_consistent(name):
    with open(os.path.join(REPO, "profiles", name)) as f:
        d = json.load(f)
    stages = d["per_stage_flops_per_env_step"]
    assert d["flops_per_env_step"] > 0 and all(v >= 0 for v in stages.values())
    # the stage split is sampled on every 10th substep; its sum stays near the mean over all substeps
    assert abs(sum(stages.values()) / d["flops_per_env_step"] - 1.0) < 0.25
