"""Benchmark scenes (models + seeded initial states / targets) shared by tests, smoke() and bench.py.

Franka: plane + the bundled Franka Panda MJCF (9 DoF, convex-hull collision geoms, self-collision on), dt = 0.01 s.
Initial joint positions and PD targets are drawn uniformly inside the joint limits with numpy PCG64 generators
(initial state: seed 0; targets of the k-th 100-step block: seed 1000 + k), sampled in float64.
"""

from __future__ import annotations

import numpy as np

from genesis_amd.model.compiler import SolverOptions, compile_model
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
