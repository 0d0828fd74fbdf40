"""meaninertia as kernel_init_meaninertia (genesis/engine/solvers/rigid/abd/misc.py:103-123) computes it.

The reference keeps a running sum and divides it by the total dof count once per entity, inside the entity loop, so
the diagonal block of entity e (of E, in order) ends up divided by D^(E - e). The checks below use that closed form
(not the loop) and the single-entity case trace(M)/D.
"""

from __future__ import annotations

import numpy as np

from genesis_amd.model.compiler import mass_matrix, meaninertia_from_mass
from genesis_amd.scenes import box_pile_model, franka_model, quadruped_model


def _closed_form(model):
    A = model.arrays
    M, *_ = mass_matrix(model, A["qpos0"])
    D = model.n_dofs
    starts, ends = A["entity_dof_start"], A["entity_dof_end"]
    E = len(starts)
    return sum(np.trace(M[s:e, s:e]) / float(D) ** (E - i) for i, (s, e) in enumerate(zip(starts, ends))), M


def test_box_pile_meaninertia_follows_the_per_entity_division():
    model = box_pile_model()
    expect, M = _closed_form(model)
    assert abs(model.meaninertia - expect) <= 1e-12 * expect
    # the plane first (no dofs), then 20 boxes: far below trace(M)/D
    assert abs(model.meaninertia - 0.01740) < 5e-5
    assert model.meaninertia < 0.1 * np.trace(M) / model.n_dofs


def test_single_robot_meaninertia_is_trace_over_dofs():
    for model in (franka_model(), quadruped_model()):
        M, *_ = mass_matrix(model, model.arrays["qpos0"])
        assert abs(model.meaninertia - np.trace(M) / model.n_dofs) <= 1e-12 * model.meaninertia


def test_entity_order_matters_like_the_reference_loop():
    # two entities of 1 dof each: ((0 + a)/2 + b)/2 for the order (a, b)
    M = np.diag([3.0, 5.0])
    assert meaninertia_from_mass(M, [0, 1], [1, 2]) == ((3.0 / 2) + 5.0) / 2
    assert meaninertia_from_mass(M, [1, 0], [2, 1]) == ((5.0 / 2) + 3.0) / 2
    assert meaninertia_from_mass(np.zeros((0, 0)), [], []) == 1.0
