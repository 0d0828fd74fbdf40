"""Joint friction loss (constraint/solver.py:580-1286 friction-loss rows, :2620-2694 their update): one row per dof with
frictionloss > 0, whose force stays within [-f, f].

CPU pins of the restatement: a hinge driven by a constant torque tau about its axis (no gravity). With tau > f the dof
slides, the row sits in its linear region and its force is exactly -f, so M acc = tau - f (M read from the restatement;
the identity holds to rounding). With tau < f the row holds the dof: the acceleration is a small fraction of tau / M
(soft constraint). The sign follows the direction of motion. The GPU tests run the same model batched through both
device paths against the FP32 restatement.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

from genesis_amd.model.compiler import SolverOptions, compile_model
from genesis_amd.model.mjcf import load_mjcf
from tests.oracle_lib import Oracle
from tests.parity import get_dev, scaled_err, slab_to_oracle

HINGE = """<mujoco model="rotor"><option timestep="0.01"/><worldbody>
  <body name="rotor" pos="0 0 1"><joint name="hinge" type="hinge" axis="0 0 1" frictionloss="{f}"/>
    <inertial pos="0.1 0 0" mass="2" diaginertia="0.01 0.02 0.03"/>
    <geom type="sphere" size="0.05" pos="0.1 0 0" contype="0" conaffinity="0"/></body>
</worldbody></mujoco>"""

FORCE = 2


def _model(tmp_path, f=0.5):
    p = os.path.join(str(tmp_path), "rotor.xml")
    with open(p, "w") as fh:
        fh.write(HINGE.format(f=f))
    return compile_model([load_mjcf(p, "rotor")], SolverOptions(dt=0.01, gravity=(0.0, 0.0, 0.0), integrator=0))


def _run(m, taus, steps, precision=64):
    n = len(taus)
    o = Oracle(m, n, precision=precision)
    o.set("ctrl_mode", np.full((n, m.n_dofs), FORCE))
    o.set("ctrl_force", np.asarray(taus, np.float64)[:, None])
    o.forward_kinematics()
    o.step(steps, 1)
    return o


def test_model_has_one_friction_row(tmp_path):
    m = _model(tmp_path)
    assert m.n_dofs == 1 and float(np.asarray(m.arrays["dof_frictionloss"]).ravel()[0]) == 0.5
    o = _run(m, [1.0], 1)
    assert int(o.get("n_con_fl")[0, 0]) == 1 and int(o.get("n_con")[0, 0]) == 1


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_sliding_dof_feels_exactly_the_friction_loss(tmp_path, sign):
    f = 0.5
    m = _model(tmp_path, f)
    taus = sign * np.array([0.8, 1.5, 3.0])
    o = _run(m, taus, 20)
    M = o.get("mass_mat")[:, 0]
    acc, vel = o.get("acc")[:, 0], o.get("vel")[:, 0]
    assert np.all(np.sign(vel) == sign)
    # linear region: the row's force is -f sign(v), so M acc = tau - f sign(v)
    assert np.allclose(M * acc, taus - sign * f, rtol=1e-9, atol=1e-12), (M * acc, taus - sign * f)
    assert np.allclose(o.get("qf_constraint")[:, 0], -sign * f, rtol=1e-9)


def test_friction_loss_holds_a_small_torque(tmp_path):
    f = 0.5
    m = _model(tmp_path, f)
    taus = np.array([0.1, -0.2, 0.45])
    o = _run(m, taus, 20)
    M = o.get("mass_mat")[:, 0]
    acc = o.get("acc")[:, 0]
    # the row opposes the torque (force within [-f, f]); only the soft constraint's compliance lets the dof creep
    assert np.all(np.abs(o.get("qf_constraint")[:, 0]) <= f + 1e-12)
    assert np.all(np.abs(M * acc) < 0.05 * np.abs(taus)), M * acc / taus


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["plain", "fused"])
def test_friction_rows_on_the_device_match_the_restatement(tmp_path, mode):
    import torch

    from genesis_amd import runtime as rt

    rt.load_library()
    m = _model(tmp_path)
    B = 256
    rng = np.random.default_rng(2)
    taus = rng.uniform(-2.0, 2.0, B)  # both regimes, both signs
    dm = rt.DeviceModel(m)
    slab = rt.StateSlab(dm, B)
    slab["ctrl_mode"].fill_(FORCE)
    slab["ctrl_force"].copy_(torch.as_tensor(taus[None, :], dtype=torch.float32))
    slab.forward_kinematics()
    slab.step(5, rt.RS_MODE_PLAIN)
    torch.cuda.synchronize()
    o = Oracle(m, B, precision=32)
    slab_to_oracle(slab, o)
    slab.step(1, rt.RS_MODE_PLAIN if mode == "plain" else rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    o.step(1, 1)
    assert np.array_equal(get_dev(slab, "n_con_fl"), o.get("n_con_fl"))
    for name in ("qpos", "vel", "qf_constraint"):
        err = scaled_err(get_dev(slab, name), o.get(name))
        assert err < 5e-5, f"{name}: {err:.3e}"
    d, r = get_dev(slab, "acc"), o.get("acc")
    scale = np.maximum(np.abs(r).max(axis=1, keepdims=True), 1.0)
    assert np.max(np.abs(d - r) / scale) < 2e-3
