"""The conjugate-gradient constraint solver on the device (constraint/solver.py:3063-3139, solver_type CG): the
per-stage step against the FP32 CPU restatement on identical input, and the fused step against the per-stage step in
each row-tile class (bit for bit up to 24 dofs; above, where the fused solver's dense row products run on the matrix
cores and the 120-dof box pile stops at the 50-iteration cap, as close to the FP64 restatement as the per-stage
step, check_fused_as_accurate_as_plain) and on equality rows."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from genesis_amd.model.specs import JOINT_REVOLUTE
from genesis_amd.scenes import (box_pile_model, closed_loops_model, franka_initial_qpos, franka_model, franka_targets,
                                quadruped_ctrl_mode, quadruped_initial_qpos, quadruped_model, quadruped_targets)
from tests.oracle_lib import Oracle
from tests.parity import (MFMA_MIN_DOFS, check_fused_as_accurate_as_plain, check_fused_matches_plain, get_dev, scaled_err,
                          slab_to_oracle)

pytestmark = pytest.mark.gpu

CG = 0
OUTPUTS = ["qpos", "vel", "acc", "qacc_ws", "links_pos", "links_quat", "mass_mat", "qf_constraint", "force",
           "n_contacts", "n_con", "errbits"]


def _rt():
    from genesis_amd import runtime

    runtime.load_library()
    return runtime


def _franka(B, warm):
    rt = _rt()
    model = franka_model(constraint_solver=CG)
    dm = rt.DeviceModel(model)
    slab = rt.StateSlab(dm, B)
    slab["qpos"].copy_(torch.as_tensor(franka_initial_qpos(model, B, 0).T, dtype=torch.float32))
    slab["ctrl_mode"].fill_(0)
    slab["ctrl_pos"].copy_(torch.as_tensor(franka_targets(model, B, 0).T, dtype=torch.float32))
    slab.forward_kinematics()
    slab.step(warm, rt.RS_MODE_PLAIN)
    torch.cuda.synchronize()
    return rt, model, dm, slab


def _quadruped(B, warm):
    rt = _rt()
    model = quadruped_model(constraint_solver=CG)
    dm = rt.DeviceModel(model)
    slab = rt.StateSlab(dm, B)
    slab["qpos"].copy_(torch.as_tensor(quadruped_initial_qpos(model, B).T, dtype=torch.float32))
    slab["ctrl_mode"].copy_(torch.as_tensor(quadruped_ctrl_mode(model, B).T, dtype=torch.int32))
    slab["ctrl_pos"].copy_(torch.as_tensor(quadruped_targets(model, B, 0).T, dtype=torch.float32))
    slab.forward_kinematics()
    slab.step(warm, rt.RS_MODE_PLAIN)
    torch.cuda.synchronize()
    return rt, model, dm, slab


def _box_pile(B, warm):
    rt = _rt()
    model = box_pile_model(constraint_solver=CG)
    dm = rt.DeviceModel(model)
    slab = rt.StateSlab(dm, B)
    slab.step(warm, rt.RS_MODE_PLAIN)
    torch.cuda.synchronize()
    return rt, model, dm, slab


def _loops(B, warm):
    rt = _rt()
    model = closed_loops_model(constraint_solver=CG)
    dm = rt.DeviceModel(model)
    slab = rt.StateSlab(dm, B)
    rng = np.random.default_rng(0)
    A = model.arrays
    q = np.tile(A["qpos0"], (B, 1))
    for j in range(model.n_joints):
        if A["joint_type"][j] == JOINT_REVOLUTE:
            q[:, A["joint_q_start"][j]] += rng.uniform(-0.3, 0.3, B)
    slab["qpos"].copy_(torch.as_tensor(q.T, dtype=torch.float32))
    slab["vel"].copy_(torch.as_tensor(rng.normal(0.0, 0.5, (model.n_dofs, B)), dtype=torch.float32))
    slab.forward_kinematics()
    slab.step(warm, rt.RS_MODE_PLAIN)
    torch.cuda.synchronize()
    return rt, model, dm, slab


def _max_abs(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b))) if a.size else 0.0


@pytest.mark.parametrize("case", ["franka", "quadruped", "box_pile", "loops"])
def test_cg_plain_step_tracks_fp32_oracle(case):
    make = {"franka": (_franka, 128, 30), "quadruped": (_quadruped, 128, 80), "box_pile": (_box_pile, 16, 60),
            "loops": (_loops, 64, 5)}[case]
    rt, model, dm, slab = make[0](make[1], make[2])
    oracle = Oracle(model, make[1], precision=32)
    slab_to_oracle(slab, oracle)
    print(f"[cg] plain vs oracle {case} B={make[1]}", flush=True)
    slab.step(1, rt.RS_MODE_PLAIN)
    torch.cuda.synchronize()
    oracle.step(1)
    assert np.array_equal(get_dev(slab, "n_con"), oracle.get("n_con"))
    assert int(get_dev(slab, "n_con").max()) > 0
    for name in ("qpos", "vel"):
        err = scaled_err(get_dev(slab, name), oracle.get(name))
        assert err < 5e-5, f"{name}: {err:.3e}"
    d, o = get_dev(slab, "acc"), oracle.get("acc")
    scale = np.maximum(np.abs(o).max(axis=1, keepdims=True), 1.0)
    assert np.max(np.abs(d - o) / scale) < 2e-3


@pytest.mark.parametrize("case", ["franka", "quadruped", "box_pile", "loops"])
def test_cg_fused_step_equals_plain_step(case):
    make = {"franka": (_franka, 256, 30), "quadruped": (_quadruped, 256, 80), "box_pile": (_box_pile, 32, 60),
            "loops": (_loops, 64, 5)}[case]
    rt, model, dm, a = make[0](make[1], make[2])
    b = rt.StateSlab(dm, a.n_envs)
    b.buffer.copy_(a.buffer)
    large = model.n_dofs >= MFMA_MIN_DOFS
    n = 1 if large else 3
    if large:
        oracle = Oracle(model, a.n_envs, precision=64)
        slab_to_oracle(a, oracle)
    print(f"[cg] fused vs plain {case} B={make[1]} steps={n} config={dm.fused_config}", flush=True)
    a.step(n, rt.RS_MODE_PLAIN)
    b.step(n, rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    assert int(get_dev(a, "n_con").max()) > 0
    if large:
        oracle.step(1)
        check_fused_as_accurate_as_plain(a, b, oracle, model)
    else:
        check_fused_matches_plain(a, b, model, OUTPUTS)
