"""Integrators Euler and implicitfast on the device (rigid_solver.py:2818-2826, forward_dynamics.py:1788-1859): the
per-stage step against the FP32 CPU restatement on identical input, and the fused step bit for bit against the
per-stage step (Franka: up to 16 dofs; quadruped: 17 to 24; box pile: the generic path, undamped free bodies)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from genesis_amd.scenes import (box_pile_model, franka_initial_qpos, franka_model, franka_targets,
                                quadruped_ctrl_mode, quadruped_initial_qpos, quadruped_model, quadruped_targets)
from tests.oracle_lib import Oracle
from tests.parity import MFMA_MIN_DOFS, check_fused_matches_plain, get_dev, scaled_err, slab_to_oracle

pytestmark = pytest.mark.gpu

EULER, IMPLICITFAST = 0, 1
OUTPUTS = ["qpos", "vel", "acc", "qacc_ws", "links_pos", "links_quat", "mass_mat", "qf_constraint", "force",
           "n_contacts", "n_con", "errbits"]
CASES = {"franka": (128, 30), "quadruped": (128, 80), "box_pile": (16, 60)}


def _slab(case, integrator, B, warm):
    from genesis_amd import runtime as rt

    rt.load_library()
    if case == "franka":
        model = franka_model(integrator=integrator)
    elif case == "quadruped":
        model = quadruped_model(integrator=integrator)
    else:
        model = box_pile_model(integrator=integrator)
    dm = rt.DeviceModel(model)
    slab = rt.StateSlab(dm, B)
    if case == "franka":
        slab["qpos"].copy_(torch.as_tensor(franka_initial_qpos(model, B, 0).T, dtype=torch.float32))
        slab["ctrl_mode"].fill_(0)
        slab["ctrl_pos"].copy_(torch.as_tensor(franka_targets(model, B, 0).T, dtype=torch.float32))
    elif case == "quadruped":
        slab["qpos"].copy_(torch.as_tensor(quadruped_initial_qpos(model, B).T, dtype=torch.float32))
        slab["ctrl_mode"].copy_(torch.as_tensor(quadruped_ctrl_mode(model, B).T, dtype=torch.int32))
        slab["ctrl_pos"].copy_(torch.as_tensor(quadruped_targets(model, B, 0).T, dtype=torch.float32))
    slab.forward_kinematics()
    slab.step(warm, rt.RS_MODE_PLAIN)
    torch.cuda.synchronize()
    return rt, model, dm, slab


@pytest.mark.parametrize("integrator", [EULER, IMPLICITFAST])
@pytest.mark.parametrize("case", list(CASES))
def test_integrator_plain_step_tracks_fp32_oracle(case, integrator):
    B, warm = CASES[case]
    rt, model, dm, slab = _slab(case, integrator, B, warm)
    oracle = Oracle(model, B, precision=32)
    slab_to_oracle(slab, oracle)
    print(f"[integrator] plain vs oracle {case} integrator={integrator} B={B}", flush=True)
    slab.step(1, rt.RS_MODE_PLAIN)
    torch.cuda.synchronize()
    oracle.step(1)
    assert np.array_equal(get_dev(slab, "n_con"), oracle.get("n_con"))
    for name in ("qpos", "vel"):
        err = scaled_err(get_dev(slab, name), oracle.get(name))
        assert err < 5e-5, f"{name}: {err:.3e}"
    d, o = get_dev(slab, "acc"), oracle.get("acc")
    scale = np.maximum(np.abs(o).max(axis=1, keepdims=True), 1.0)
    assert np.max(np.abs(d - o) / scale) < 2e-3


@pytest.mark.parametrize("integrator", [EULER, IMPLICITFAST])
@pytest.mark.parametrize("case", list(CASES))
def test_integrator_fused_step_equals_plain_step(case, integrator):
    B, warm = CASES[case]
    rt, model, dm, a = _slab(case, integrator, 2 * B, warm)
    b = rt.StateSlab(dm, a.n_envs)
    b.buffer.copy_(a.buffer)
    print(f"[integrator] fused vs plain {case} integrator={integrator} config={dm.fused_config}", flush=True)
    n = 3 if model.n_dofs < MFMA_MIN_DOFS else 1
    a.step(n, rt.RS_MODE_PLAIN)
    b.step(n, rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    check_fused_matches_plain(a, b, model, OUTPUTS)
