"""The conjugate-gradient constraint solver (constraint/solver.py:3063-3139 with solver_type CG: the gradient
preconditioned by the mass matrix, :2737-2747, and the Polak-Ribiere direction, :2865-2884) in the CPU restatement.

Pins (FP64):
- the constrained acceleration minimises a strictly convex cost (the Gauss term is positive definite), so CG and
  Newton started from the same state must reach the same acceleration, with the contact, friction and limit rows
  of the Franka arm, the box pile and the quadruped;
- on a model whose rows are all equality rows (the closed-loop mechanisms: no contacts, no limits) the cost is one
  quadratic and the line search exact, so preconditioned CG reaches the minimum in at most n_dofs steps (the
  preconditioned Hessian is the identity plus a term of rank n_rows < n_dofs); the iteration count then includes
  at most one more iteration that finds nothing left to improve (Newton: one step, one or two iterations).
  Steepest descent (beta forced to zero) or a wrong sign or operand in beta needs many more iterations there;
- with no constraint rows the solve returns the unconstrained acceleration unchanged, without iterating.
"""

from __future__ import annotations

import numpy as np
import pytest

from genesis_amd.model.specs import JOINT_REVOLUTE
from genesis_amd.scenes import (box_pile_model, closed_loops_model, franka_initial_qpos, franka_model,
                                quadruped_initial_qpos, quadruped_model)
from tests.oracle_lib import Oracle

CG, NEWTON = 0, 1
# fields that carry the state from one step to the next: configuration, warm start, controls, and the collision
# caches (sweep-and-prune order and the MPR warm-start normals)
CARRY = ("qpos", "vel", "qacc_ws", "is_warmstart", "ctrl_mode", "ctrl_pos", "sort_value", "sort_i_g", "sort_is_max",
         "cache_normal")


def _same_state_solve(make, warm, prep=None, B=4):
    """Warm up with Newton, then take one more step with each solver from the same carried state."""
    opts = dict(tolerance=1e-12, iterations=1000)
    mn = make(**opts)
    on = Oracle(mn, B, precision=64)
    if prep is not None:
        prep(mn, on)
    on.forward_kinematics()
    on.step(warm)
    oc = Oracle(make(constraint_solver=CG, **opts), B, precision=64)
    for f in CARRY:
        oc.set(f, on.get(f))
    oc.forward_kinematics()
    on.step(1)
    oc.step(1)
    return on, oc


@pytest.mark.parametrize("case", ["franka", "box_pile", "quadruped"])
def test_cg_and_newton_reach_the_same_acceleration(case):
    if case == "franka":
        on, oc = _same_state_solve(franka_model, 40, lambda m, o: o.set("qpos", franka_initial_qpos(m, 4, 0)))
    elif case == "box_pile":
        on, oc = _same_state_solve(box_pile_model, 60)
    else:
        on, oc = _same_state_solve(quadruped_model, 80, lambda m, o: o.set("qpos", quadruped_initial_qpos(m, 4)))
    assert np.array_equal(on.get("n_con"), oc.get("n_con"))
    assert int(on.get("n_con").max()) > 0
    a_n, a_c = on.get("acc"), oc.get("acc")
    scale = np.maximum(np.abs(a_n).max(axis=1), 1.0)
    err = np.abs(a_n - a_c).max(axis=1) / scale
    assert err.max() < 1e-5, err
    # CG needs more iterations than Newton wherever the rows are not trivial
    assert oc.get("n_iter").sum() > on.get("n_iter").sum()


def _loops(solver, B=8, seed=0):
    model = closed_loops_model(constraint_solver=solver, tolerance=1e-12, iterations=200)
    o = Oracle(model, B, precision=64)
    rng = np.random.default_rng(seed)
    A = model.arrays
    q = np.tile(A["qpos0"], (B, 1))
    for j in range(model.n_joints):
        if A["joint_type"][j] == JOINT_REVOLUTE:
            q[:, A["joint_q_start"][j]] += rng.uniform(-0.3, 0.3, B)
    o.set("qpos", q)
    o.set("vel", rng.normal(0.0, 0.5, (B, model.n_dofs)))
    o.forward_kinematics()
    return model, o


def test_cg_ends_within_n_dofs_iterations_on_a_quadratic():
    model, oc = _loops(CG)
    _, on = _loops(NEWTON)
    for _ in range(3):
        oc.step(1)
        on.step(1)
        it = oc.get("n_iter").ravel()
        assert int(oc.get("n_con").min()) == 9  # connect (3 rows) + weld (6 rows), nothing else
        assert it.max() <= model.n_dofs + 1, it
        assert on.get("n_iter").max() <= 2
        a_n, a_c = on.get("acc"), oc.get("acc")
        scale = np.maximum(np.abs(a_n).max(axis=1), 1.0)
        assert (np.abs(a_n - a_c).max(axis=1) / scale).max() < 1e-7
        for f in ("qpos", "vel", "qacc_ws"):  # keep the two rollouts on one trajectory
            oc.set(f, on.get(f))
        oc.forward_kinematics()


def test_cg_without_rows_returns_the_unconstrained_acceleration():
    model = box_pile_model(constraint_solver=CG, enable_collision=False)  # free bodies: no limits, no friction
    o = Oracle(model, 4, precision=64)
    o.forward_kinematics()
    o.step(1)
    assert int(o.get("n_con").max()) == 0
    assert int(o.get("n_iter").max()) == 0
    np.testing.assert_array_equal(o.get("acc"), o.get("acc_smooth"))


def test_rigid_options_accept_both_solvers():
    import genesis_amd as gs

    for solver in (gs.constraint_solver.CG, gs.constraint_solver.Newton):
        gs.options.RigidOptions(constraint_solver=solver).validate()
    with pytest.raises(ValueError):
        gs.options.RigidOptions(constraint_solver=7).validate()
