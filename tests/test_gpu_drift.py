"""Trajectory drift of the FP32 device step against the FP64 CPU restatement (the stand-in for Genesis CPU at
precision='64'), over fixed horizons, for every workload.

Stated tolerances (set from the FP32 CPU restatement's own drift against FP64, measured on the same seeds):
* Franka x64, 1000 steps: median over envs of max |q32 - q64| <= 1e-4 rad, 90th percentile <= 5e-3 rad; the worst
  env within twice the FP32 CPU restatement's own worst drift (and within 1e-2 rad if that is smaller), and at most
  twice as many envs as the FP32 restatement (plus one) beyond 1e-2 rad. A few envs start in deep self-penetration and
  move chaotically: the FP32 restatement itself ends up to ~0.5 rad from FP64 on one or two of the 64, so a flat 1e-2
  rad bound on the worst env would fail for any FP32 implementation;
* quadruped x64, 1000 steps: base position/orientation drift <= 1e-4 for every env;
* box pile x32: per-box position drift <= 1e-4 m over the first 50 steps; after 500 steps (settled, chaotic order of
  contacts) the mean resting height agrees to 1e-3 m.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from genesis_amd import scenes
from tests.oracle_lib import Oracle
from tests.parity import get_dev

pytestmark = pytest.mark.gpu


def _rt():
    from genesis_amd import runtime

    runtime.load_library()
    return runtime


def _pair(model, B, init_slab, init_oracle):
    rt = _rt()
    dm = rt.DeviceModel(model)
    slab = rt.StateSlab(dm, B)
    init_slab(slab)
    slab.forward_kinematics()
    o = Oracle(model, B, precision=64)
    init_oracle(o)
    o.forward_kinematics()
    return rt, slab, o


def test_franka_drift_vs_fp64():
    m = scenes.franka_model()
    B = 64
    q0, t0 = scenes.franka_initial_qpos(m, B), scenes.franka_targets(m, B, 0)

    def init_slab(s):
        s["qpos"].copy_(torch.as_tensor(q0.T, dtype=torch.float32))
        s["ctrl_mode"].fill_(0)
        s["ctrl_pos"].copy_(torch.as_tensor(t0.T, dtype=torch.float32))

    def init_oracle(o):
        o.set("qpos", q0.astype(np.float32).astype(np.float64))
        o.set("ctrl_mode", np.zeros((B, m.n_dofs)))
        o.set("ctrl_pos", t0.astype(np.float32).astype(np.float64))

    rt, slab, o = _pair(m, B, init_slab, init_oracle)
    o32 = Oracle(m, B, precision=32)
    init_oracle(o32)
    o32.forward_kinematics()
    slab.step(1000, rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    o.step(1000, 16)
    o32.step(1000, 16)
    q64 = o.get("qpos")
    d = np.abs(get_dev(slab, "qpos") - q64).max(axis=1)
    d32 = np.abs(o32.get("qpos") - q64).max(axis=1)
    print(f"[drift] franka 1000 steps: median {np.median(d):.2e} p90 {np.percentile(d, 90):.2e} max {d.max():.2e}; "
          f"FP32 restatement: median {np.median(d32):.2e} max {d32.max():.2e}, "
          f"envs > 1e-2 rad: device {(d > 1e-2).sum()} restatement {(d32 > 1e-2).sum()}", flush=True)
    assert np.median(d) <= 1e-4
    assert np.percentile(d, 90) <= 5e-3
    assert d.max() <= 2.0 * max(d32.max(), 1e-2)
    assert (d > 1e-2).sum() <= 2 * (d32 > 1e-2).sum() + 1


def test_quadruped_drift_vs_fp64():
    m = scenes.quadruped_model()
    B = 64
    q0, mode, t0 = (scenes.quadruped_initial_qpos(m, B), scenes.quadruped_ctrl_mode(m, B),
                    scenes.quadruped_targets(m, B, 0))

    def init_slab(s):
        s["qpos"].copy_(torch.as_tensor(q0.T, dtype=torch.float32))
        s["ctrl_mode"].copy_(torch.as_tensor(mode.T, dtype=torch.int32))
        s["ctrl_pos"].copy_(torch.as_tensor(t0.T, dtype=torch.float32))

    def init_oracle(o):
        o.set("qpos", q0.astype(np.float32).astype(np.float64))
        o.set("ctrl_mode", mode)
        o.set("ctrl_pos", t0.astype(np.float32).astype(np.float64))

    rt, slab, o = _pair(m, B, init_slab, init_oracle)
    slab.step(1000, rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    o.step(1000, 16)
    d = np.abs(get_dev(slab, "qpos")[:, :7] - o.get("qpos")[:, :7]).max(axis=1)
    print(f"[drift] quadruped base, 1000 steps: median {np.median(d):.2e} max {d.max():.2e}")
    assert d.max() <= 1e-4


def test_box_pile_drift_vs_fp64():
    m = scenes.box_pile_model()
    B = 32
    rt, slab, o = _pair(m, B, lambda s: None, lambda o: None)
    pos = lambda q: q.reshape(B, -1, 7)[:, :, :3]  # noqa: E731
    slab.step(50, rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    o.step(50, 16)
    d = np.linalg.norm(pos(get_dev(slab, "qpos")) - pos(o.get("qpos")), axis=2)
    print(f"[drift] box pile 50 steps: max {d.max():.2e}")
    assert d.max() <= 1e-4
    slab.step(450, rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    o.step(450, 16)
    z32, z64 = pos(get_dev(slab, "qpos"))[:, :, 2].mean(), pos(o.get("qpos"))[:, :, 2].mean()
    print(f"[drift] box pile 500 steps: mean rest height {z32:.5f} vs {z64:.5f}")
    assert abs(z32 - z64) <= 1e-3
