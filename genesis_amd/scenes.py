"""Benchmark scenes (models + seeded initial states / targets) shared by tests, smoke() and bench.py.

Franka: plane + the bundled Franka Panda MJCF (9 DoF, convex-hull collision geoms, self-collision on), dt = 0.01 s.
Initial joint positions and PD targets are drawn uniformly inside the joint limits with numpy PCG64 generators
(initial state: seed 0; targets of the k-th 100-step block: seed 1000 + k), sampled in float64.

Box pile: 20 free boxes, sides U[0.05, 0.15] m, density 1000, friction 0.5, dropped from U[0.2, 1.5] m with random
orientations (PCG64 seed 2) onto a plane, dt = 0.01 s.

Quadruped: plane + the bundled quadruped MJCF (free base, box torso, 4 legs x 3 hinges, capsule legs, sphere feet,
18 DoF), dt = 0.005 s, starting in the standing pose; PD targets = standing pose + U[-0.3, 0.3] per joint (PCG64
seed 3 for the first 100-step block, 3 + 1000 k for block k).
"""

from __future__ import annotations

import os

import numpy as np

from genesis_amd.model.compiler import SolverOptions, compile_model
from genesis_amd.model.mjcf import load_mjcf
from genesis_amd.model.primitives import box_entity, plane_entity
from genesis_amd.morphs import load_mjcf_or_cached

FRANKA_XML = "xml/franka_emika_panda/panda.xml"


def franka_model(**opts):
    return compile_model([plane_entity(), load_mjcf_or_cached(FRANKA_XML, "franka")], SolverOptions(dt=0.01, **opts))


def joint_limits(model):
    A = model.arrays
    lim = A["dof_limit"].copy()
    return lim[:, 0], lim[:, 1]


def franka_initial_qpos(model, n_envs, seed=0):
    lo, hi = joint_limits(model)
    rng = np.random.Generator(np.random.PCG64(seed))
    return lo[None] + (hi - lo)[None] * rng.random((n_envs, len(lo)))


def franka_targets(model, n_envs, block):
    lo, hi = joint_limits(model)
    rng = np.random.Generator(np.random.PCG64(1000 + block))
    return lo[None] + (hi - lo)[None] * rng.random((n_envs, len(lo)))


def box_pile_model(n_boxes=20, seed=2, **opts):
    """20 free boxes (sides U[0.05, 0.15], density 1000, friction 0.5) dropped on a plane."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ents = [plane_entity(friction=0.5)]
    for i in range(n_boxes):
        size = rng.uniform(0.05, 0.15, 3)
        pos = np.array([*rng.uniform(-0.25, 0.25, 2), rng.uniform(0.2, 1.5)])
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        ents.append(box_entity(size, pos, q, rho=1000.0, friction=0.5, name=f"box{i}"))
    return compile_model(ents, SolverOptions(dt=0.01, **opts))


QUADRUPED_XML = os.path.join(os.path.dirname(os.path.abspath(__file__)), "assets", "xml", "quadruped.xml")
QUADRUPED_STAND = np.array([0.0, 0.0, 0.0, 0.0, 0.7, 0.7, 0.7, 0.7, -1.4, -1.4, -1.4, -1.4])  # abd x4, hip x4, knee x4
QUADRUPED_BASE_Z = 0.36


CLOSED_LOOPS_XML = os.path.join(os.path.dirname(os.path.abspath(__file__)), "assets", "xml", "closed_loops.xml")


def closed_loops_model(**opts):
    """A four-bar linkage closed by a connect equality and a free box welded to a pendulum (no collisions), dt 0.005."""
    return compile_model([load_mjcf(CLOSED_LOOPS_XML, "loops")], SolverOptions(dt=0.005, **opts))


def quadruped_model(**opts):
    return compile_model([plane_entity(), load_mjcf(QUADRUPED_XML, "quadruped")], SolverOptions(dt=0.005, **opts))


def quadruped_initial_qpos(model, n_envs):
    q = np.zeros((n_envs, model.n_qs))
    q[:, 2] = QUADRUPED_BASE_Z
    q[:, 3] = 1.0
    q[:, 7:] = QUADRUPED_STAND[None]
    return q


def quadruped_ctrl_mode(model, n_envs):
    """FORCE (unactuated, zero force) on the six base dofs, POSITION on the twelve joints."""
    mode = np.zeros((n_envs, model.n_dofs))
    mode[:, :6] = 2
    return mode


def quadruped_targets(model, n_envs, block):
    rng = np.random.Generator(np.random.PCG64(3 + 1000 * block))
    t = np.zeros((n_envs, model.n_dofs))
    t[:, 6:] = QUADRUPED_STAND[None] + rng.uniform(-0.3, 0.3, (n_envs, 12))
    return t
