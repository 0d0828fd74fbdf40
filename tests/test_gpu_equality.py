"""Connect and weld equality rows on the device (constraint/solver.py:689-797, :984-1173): the per-stage rows
against the FP32 CPU restatement on identical input, the fused step bit for bit against the per-stage step, and a
rollout that keeps the anchors together."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from genesis_amd.model import transforms as tf
from genesis_amd.model.specs import EQ_CONNECT, JOINT_REVOLUTE
from genesis_amd.scenes import closed_loops_model
from tests.oracle_lib import Oracle
from tests.parity import get_dev, scaled_err, slab_to_oracle

pytestmark = pytest.mark.gpu

OUTPUTS = ["qpos", "vel", "acc", "qacc_ws", "links_pos", "links_quat", "qf_constraint", "force", "n_con", "errbits"]


def _rt():
    from genesis_amd import runtime

    runtime.load_library()
    return runtime


def _slab(B, seed=0, layout=None):
    rt = _rt()
    model = closed_loops_model()
    dm = rt.DeviceModel(model, rt.LAYOUT_ENV_RECORDS if layout is None else layout)
    slab = rt.StateSlab(dm, B)
    rng = np.random.default_rng(seed)
    A = model.arrays
    q = np.tile(A["qpos0"], (B, 1))
    for j in range(model.n_joints):
        if A["joint_type"][j] == JOINT_REVOLUTE:
            q[:, A["joint_q_start"][j]] += rng.uniform(-0.3, 0.3, B)
    slab["qpos"].copy_(torch.as_tensor(q.T, dtype=torch.float32))
    slab["vel"].copy_(torch.as_tensor(rng.normal(0.0, 0.5, (model.n_dofs, B)), dtype=torch.float32))
    slab.forward_kinematics()
    torch.cuda.synchronize()
    return rt, model, dm, slab


def test_equality_rows_match_the_restatement():
    rt, model, dm, slab = _slab(64, layout=_rt().LAYOUT_PLANES)
    oracle = Oracle(model, 64, precision=32)
    print("[equality] stage parity B=64", flush=True)
    for stage in ("crba", "factor", "rne", "constraints"):
        slab_to_oracle(slab, oracle)
        slab.stage(stage)
        torch.cuda.synchronize()
        oracle.stage(rt.STAGE_ID[stage])
    n_d, n_o = get_dev(slab, "n_con"), oracle.get("n_con")
    assert np.array_equal(n_d, n_o) and int(n_d.min()) >= 9
    assert np.array_equal(get_dev(slab, "n_con_eq"), oracle.get("n_con_eq"))
    ne = int(get_dev(slab, "n_con_eq")[0, 0])
    D = model.n_dofs
    for name, width in (("jac", D), ("aref", 1), ("diag", 1), ("efc_D", 1)):
        d, o = get_dev(slab, name)[:, : ne * width], oracle.get(name)[:, : ne * width]
        assert scaled_err(d, o) < 1e-4, name


@pytest.mark.parametrize("layout", ["records", "planes"])
def test_fused_equals_plain_with_equalities(layout):
    rt = _rt()
    _, model, dm, a = _slab(128, seed=1, layout=rt.LAYOUT_PLANES if layout == "planes" else rt.LAYOUT_ENV_RECORDS)
    b = rt.StateSlab(dm, a.n_envs)
    b.buffer.copy_(a.buffer)
    print(f"[equality] fused vs plain layout={layout} config={dm.fused_config}", flush=True)
    a.step(20, rt.RS_MODE_PLAIN)
    b.step(20, rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    for name in OUTPUTS:
        da, db = np.asarray(get_dev(a, name), np.float64), np.asarray(get_dev(b, name), np.float64)
        assert np.max(np.abs(da - db)) == 0.0, name


def test_rollout_keeps_anchors_together_and_tracks_the_restatement():
    rt, model, dm, slab = _slab(32, seed=2)
    oracle = Oracle(model, 32, precision=32)
    slab_to_oracle(slab, oracle)
    print("[equality] rollout 200 substeps", flush=True)
    slab.step(200, rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    oracle.step(200)
    q_d, q_o = get_dev(slab, "qpos"), oracle.get("qpos")
    assert np.median(np.abs(q_d - q_o).max(axis=1)) < 1e-3
    A = model.arrays
    lpos, lquat = get_dev(slab, "links_pos"), get_dev(slab, "links_quat")
    for b in range(slab.n_envs):
        P, Q = lpos[b].reshape(-1, 3).astype(np.float64), lquat[b].reshape(-1, 4).astype(np.float64)
        for k in range(model.n_equalities):
            l1, l2, d = int(A["eq_obj1"][k]), int(A["eq_obj2"][k]), A["eq_data"][k]
            if A["eq_type"][k] == EQ_CONNECT:
                gap = tf.transform_point(d[0:3], P[l1], Q[l1]) - tf.transform_point(d[3:6], P[l2], Q[l2])
            else:
                gap = tf.transform_point(d[3:6], P[l1], Q[l1]) - tf.transform_point(d[0:3], P[l2], Q[l2])
            assert np.linalg.norm(gap) < 5e-3, (b, k, np.linalg.norm(gap))
