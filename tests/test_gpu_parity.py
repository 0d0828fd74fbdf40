"""Device kernels vs the CPU restatement (FP32) on identical input state, stage by stage, and over short rollouts.

Tolerances: continuous fields agree to 5e-5 of the field's magnitude (the reference's own single-precision tolerance,
tests/conftest.py:78 TOL_SINGLE); integer outputs (broadphase pair lists, contact geom pairs, row counts) are equal.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from genesis_amd.scenes import box_pile_model, franka_initial_qpos, franka_model, franka_targets
from tests.oracle_lib import Oracle
from tests.parity import get_dev, scaled_err, slab_to_oracle

pytestmark = pytest.mark.gpu

TOL = 5e-5


def _rt():
    from genesis_amd import runtime

    runtime.load_library()
    return runtime


def _franka_slab(B, steps=0, seed=0, layout=None):
    rt = _rt()
    model = franka_model()
    dm = rt.DeviceModel(model, rt.LAYOUT_ENV_RECORDS if layout is None else layout)
    slab = rt.StateSlab(dm, B)
    slab["qpos"].copy_(torch.as_tensor(franka_initial_qpos(model, B, seed).T, dtype=torch.float32))
    slab["ctrl_mode"].fill_(0)
    slab["ctrl_pos"].copy_(torch.as_tensor(franka_targets(model, B, 0).T, dtype=torch.float32))
    slab.forward_kinematics()
    if steps:
        slab.step(steps, rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    return model, dm, slab


def _check(slab, oracle, name, tol=TOL, rows=None):
    dev = get_dev(slab, name)
    ref = oracle.get(name if name != "errbits" else "errno")
    if rows is not None:
        dev, ref = dev[:, :rows], ref[:, :rows]
    err = scaled_err(dev, ref)
    assert err < tol, f"{name}: scaled error {err:.3e}"
    return err


def _masked(slab, oracle, name, count_name, width):
    """Compare only the first n[b] elements of a per-env list field."""
    n = get_dev(slab, count_name)[:, 0].astype(int)
    dev, ref = get_dev(slab, name), oracle.get(name)
    mask = np.zeros_like(dev, dtype=bool)
    for b, k in enumerate(n):
        mask[b, : k * width] = True
    return dev[mask], ref[mask]


@pytest.mark.parametrize("warm_steps,layout", [(0, "planes"), (20, "planes"), (20, "records")])
def test_franka_stage_parity(warm_steps, layout):
    rt = _rt()
    B = 128
    model, dm, slab = _franka_slab(B, warm_steps,
                                   layout=rt.LAYOUT_PLANES if layout == "planes" else rt.LAYOUT_ENV_RECORDS)
    oracle = Oracle(model, B, precision=32)
    for stage, outputs in [
        ("crba", ["mass_mat"]),
        ("factor", ["mass_L", "mass_D_inv"]),
        ("rne", ["qf_applied", "qf_passive", "qf_bias", "force", "acc_smooth", "cfrc_vel", "cfrc_ang"]),
        ("aabb", ["aabb_min", "aabb_max"]),
        ("broadphase", []),
        ("narrowphase", []),
        ("constraints", []),
        ("solve", []),
        ("integrate", ["qpos", "vel", "links_pos", "links_quat", "cd_vel", "cd_ang", "cdof_vel"]),
    ]:
        slab_to_oracle(slab, oracle)
        slab.stage(stage)
        torch.cuda.synchronize()
        oracle.stage(rt.STAGE_ID[stage])
        print(f"[parity] warm={warm_steps} stage={stage}", flush=True)
        for name in outputs:
            _check(slab, oracle, name)
        if stage == "broadphase":
            nb_d, nb_o = get_dev(slab, "n_broad")[:, 0], oracle.get("n_broad")[:, 0]
            assert np.array_equal(nb_d, nb_o), "broadphase pair counts differ"
            d, o = _masked(slab, oracle, "broad_pairs", "n_broad", 2)
            assert np.array_equal(d, o), "broadphase pair lists differ"
        if stage == "narrowphase":
            nc_d, nc_o = get_dev(slab, "n_contacts")[:, 0], oracle.get("n_contacts")[:, 0]
            assert np.array_equal(nc_d, nc_o), f"contact counts differ: {np.nonzero(nc_d != nc_o)}"
            for f in ("contact_geom_a", "contact_geom_b"):
                d, o = _masked(slab, oracle, f, "n_contacts", 1)
                assert np.array_equal(d, o), f"{f} differs"
            for f in ("contact_pos", "contact_normal"):
                d, o = _masked(slab, oracle, f, "n_contacts", 3)
                assert scaled_err(d, o) < 1e-4, f
            d, o = _masked(slab, oracle, "contact_penetration", "n_contacts", 1)
            if d.size:
                assert np.max(np.abs(d - o)) < 5e-5, "contact_penetration"
        if stage == "constraints":
            assert np.array_equal(get_dev(slab, "n_con"), oracle.get("n_con"))
            d, o = _masked(slab, oracle, "aref", "n_con", 1)
            assert scaled_err(d, o) < 1e-4
            d, o = _masked(slab, oracle, "efc_D", "n_con", 1)
            assert scaled_err(d, o) < 1e-4
            D = model.n_dofs
            dj, oj = _masked(slab, oracle, "jac", "n_con", D)
            assert scaled_err(dj, oj) < 1e-4
        if stage == "solve":
            # the constrained acceleration is the minimiser of a strictly convex problem: both sides must agree
            d, o = get_dev(slab, "acc"), oracle.get("acc")
            scale = np.maximum(np.abs(o).max(axis=1, keepdims=True), 1.0)
            err = np.max(np.abs(d - o) / scale)
            assert err < 2e-3, f"acc after solve: {err}"


def test_franka_rollout_close_to_fp32_oracle():
    """50 substeps from the same state: device FP32 and CPU FP32 trajectories stay close (smooth PD motion)."""
    rt = _rt()
    B = 64
    model, dm, slab = _franka_slab(B)
    oracle = Oracle(model, B, precision=32)
    slab_to_oracle(slab, oracle)
    slab.step(50, rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    oracle.step(50)
    dq = np.abs(get_dev(slab, "qpos") - oracle.get("qpos"))
    # chaotic contacts make a few envs diverge; the bulk must track closely
    assert np.median(dq.max(axis=1)) < 1e-3, np.median(dq.max(axis=1))


def test_plain_and_fused_paths_agree():
    rt = _rt()
    B = 64
    model, dm, a = _franka_slab(B)
    b = rt.StateSlab(dm, B)
    b.buffer.copy_(a.buffer)
    a.step(10, rt.RS_MODE_PLAIN)
    b.step(10, rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    for name in ("qpos", "vel", "acc"):
        assert scaled_err(get_dev(a, name), get_dev(b, name)) < 1e-4, name


def test_box_pile_stage_parity():
    rt = _rt()
    B = 32
    model = box_pile_model()
    dm = rt.DeviceModel(model)
    slab = rt.StateSlab(dm, B)
    slab.step(60, rt.RS_MODE_FUSED)  # boxes land and pile
    torch.cuda.synchronize()
    oracle = Oracle(model, B, precision=32)
    for stage in ("crba", "factor", "rne", "aabb", "broadphase", "narrowphase", "constraints", "solve"):
        slab_to_oracle(slab, oracle)
        slab.stage(stage)
        torch.cuda.synchronize()
        oracle.stage(rt.STAGE_ID[stage])
        print(f"[parity] box pile stage={stage}", flush=True)
        if stage == "crba":
            _check(slab, oracle, "mass_mat")
        if stage == "narrowphase":
            nc_d, nc_o = get_dev(slab, "n_contacts")[:, 0], oracle.get("n_contacts")[:, 0]
            assert np.array_equal(nc_d, nc_o)
        if stage == "solve":
            d, o = get_dev(slab, "acc"), oracle.get("acc")
            scale = np.maximum(np.abs(o).max(axis=1, keepdims=True), 1.0)
            assert np.max(np.abs(d - o) / scale) < 2e-3


def test_graph_replay_matches_eager():
    rt = _rt()
    B = 64
    model, dm, a = _franka_slab(B)
    b = rt.StateSlab(dm, B)
    b.buffer.copy_(a.buffer)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        g = rt.StepGraph(b, 5, rt.RS_MODE_FUSED, stream)
        for _ in range(3):
            g.launch()
            stream.synchronize()
    a.step(15, rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    for name in ("qpos", "vel"):
        assert scaled_err(get_dev(b, name), get_dev(a, name)) < 1e-6, name
