"""A URDF-loaded robot (genesis_amd/model/urdf.py: floating base, prismatic / revolute / continuous joints, merged
fixed links, box and sphere collisions) on the device: after it has fallen onto a plane and its links touch the
ground, one fused substep equals one per-stage substep bit for bit, and both follow the FP32 restatement (contact and
row counts equal, qpos/vel within 5e-5 of their magnitude)."""

from __future__ import annotations

import os

import numpy as np
import pytest
import torch

from genesis_amd.model.compiler import SolverOptions, compile_model
from genesis_amd.model.primitives import plane_entity
from genesis_amd.model.urdf import load_urdf
from tests.oracle_lib import Oracle
from tests.parity import check_fused_matches_plain, get_dev, scaled_err, slab_to_oracle
from tests.test_urdf import ARM

pytestmark = pytest.mark.gpu


def test_urdf_arm_on_the_device(tmp_path):
    from genesis_amd import runtime as rt

    rt.load_library()
    path = os.path.join(str(tmp_path), "arm.urdf")
    with open(path, "w") as f:
        f.write(ARM)
    model = compile_model([plane_entity(), load_urdf(path, pos=(0.0, 0.0, 0.3))], SolverOptions())
    B = 64
    rng = np.random.default_rng(2)
    dm = rt.DeviceModel(model)
    a = rt.StateSlab(dm, B)
    q = np.tile(model.arrays["qpos0"], (B, 1))
    q[:, 7:] = rng.uniform(-0.5, 0.5, (B, 3)) * [0.3, 1.0, 1.0]
    a["qpos"].copy_(torch.as_tensor(q.T, dtype=torch.float32))
    a["ctrl_mode"].fill_(0)
    a["ctrl_pos"].copy_(torch.as_tensor(np.tile([0, 0, 0, 0, 0, 0, 0.1, 0.5, -0.8], (B, 1)).T, dtype=torch.float32))
    a.forward_kinematics()
    a.step(60, rt.RS_MODE_PLAIN)
    torch.cuda.synchronize()
    b = rt.StateSlab(dm, B)
    b.buffer.copy_(a.buffer)
    oracle = Oracle(model, B, precision=32)
    slab_to_oracle(a, oracle)
    print(f"[urdf] arm x{B}: {model.n_dofs} dofs, {model.n_links} links, contacts "
          f"{int(get_dev(a, 'n_contacts').sum())}", flush=True)
    a.step(1, rt.RS_MODE_PLAIN)
    b.step(1, rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    oracle.step(1)
    assert int(get_dev(a, "n_contacts").max()) > 0
    check_fused_matches_plain(a, b, model, ("qpos", "vel", "acc", "n_contacts", "n_con", "errbits"))
    assert np.array_equal(get_dev(a, "n_contacts"), oracle.get("n_contacts"))
    assert np.array_equal(get_dev(a, "n_con"), oracle.get("n_con"))
    for name in ("qpos", "vel"):
        err = scaled_err(get_dev(a, name), oracle.get(name))
        assert err < 5e-5, f"{name}: {err:.3e}"
