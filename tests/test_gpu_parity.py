"""Device kernels vs the CPU restatement (FP32) on identical input state, stage by stage, and over short rollouts.

Tolerances: continuous fields agree to 5e-5 of the field's magnitude (the reference's own single-precision tolerance,
tests/conftest.py:78 TOL_SINGLE); integer outputs (broadphase pair lists, contact geom pairs, row counts) are equal.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from genesis_amd.scenes import (box_pile_model, franka_initial_qpos, franka_model, franka_targets,
                                quadruped_ctrl_mode, quadruped_initial_qpos, quadruped_model, quadruped_targets)
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


# constraint forces (efc_force per row, qf_constraint per dof, contact_force per contact) agree to this fraction of the
# environment's largest value: the solver stops at tolerance * meaninertia * max(1, n_dofs) (constraint/solver.py:
# 3063-3139), and the FP32 restatement itself stays within 3e-4 of FP64 on the same states
FORCE_TOL = 1e-3


def _per_env_scaled(dev, ref, counts, width):
    """max over envs of max|dev - ref| / max|ref| over the env's first counts[b] * width entries."""
    worst = 0.0
    for b, k in enumerate(counts):
        n = int(k) * width
        if n == 0:
            continue
        scale = max(float(np.abs(ref[b, :n]).max()), 1e-6)
        worst = max(worst, float(np.abs(dev[b, :n] - ref[b, :n]).max()) / scale)
    return worst


def _check_forces(slab, oracle, label):
    n_con = get_dev(slab, "n_con")[:, 0].astype(int)
    n_ct = get_dev(slab, "n_contacts")[:, 0].astype(int)
    D = slab.dmodel.model.n_dofs
    errs = {
        "efc_force": _per_env_scaled(get_dev(slab, "efc_force"), oracle.get("efc_force"), n_con, 1),
        "qf_constraint": _per_env_scaled(get_dev(slab, "qf_constraint"), oracle.get("qf_constraint"),
                                         np.full(len(n_con), D), 1),
        "contact_force": _per_env_scaled(get_dev(slab, "contact_force"), oracle.get("contact_force"), n_ct, 3),
    }
    print(f"[parity] {label} forces: " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    for k, v in errs.items():
        assert v <= FORCE_TOL, f"{label}: {k} per-env scaled error {v:.3e}"


STAGE_OUTPUTS = [
    ("crba", ["mass_mat"]),
    ("factor", ["mass_L", "mass_D_inv"]),
    ("rne", ["qf_applied", "qf_passive", "qf_bias", "force", "acc_smooth", "cfrc_vel", "cfrc_ang"]),
    ("aabb", ["aabb_min", "aabb_max"]),
    ("broadphase", []),
    ("narrowphase", []),
    ("constraints", []),
    ("solve", []),
    ("integrate", ["qpos", "vel", "links_pos", "links_quat", "cd_vel", "cd_ang", "cdof_vel"]),
]


def _stage_parity(rt, model, slab, label):
    """Every plain stage on the device against the same stage of the FP32 restatement, from identical input state."""
    oracle = Oracle(model, slab.n_envs, precision=32)
    D = model.n_dofs
    for stage, outputs in STAGE_OUTPUTS:
        slab_to_oracle(slab, oracle)
        slab.stage(stage)
        torch.cuda.synchronize()
        oracle.stage(rt.STAGE_ID[stage])
        print(f"[parity] {label} stage={stage}", flush=True)
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
            for f in ("contact_geom_a", "contact_geom_b", "contact_link_a", "contact_link_b", "contact_pair"):
                d, o = _masked(slab, oracle, f, "n_contacts", 1)
                assert np.array_equal(d, o), f"{f} differs"
            for f in ("contact_pos", "contact_normal"):
                d, o = _masked(slab, oracle, f, "n_contacts", 3)
                assert scaled_err(d, o) < 1e-4, f
            d, o = _masked(slab, oracle, "contact_penetration", "n_contacts", 1)
            if d.size:
                assert np.max(np.abs(d - o)) < 5e-5, "contact_penetration"
            d, o = _masked(slab, oracle, "contact_friction", "n_contacts", 1)
            assert np.array_equal(d.astype(np.float32), o.astype(np.float32)), "contact_friction"
        if stage == "constraints":
            for f in ("n_con", "n_con_eq", "n_con_fl"):
                assert np.array_equal(get_dev(slab, f), oracle.get(f)), f
            for f in ("aref", "efc_D", "diag"):
                d, o = _masked(slab, oracle, f, "n_con", 1)
                assert scaled_err(d, o) < 1e-4, f
            dj, oj = _masked(slab, oracle, "jac", "n_con", D)
            assert scaled_err(dj, oj) < 1e-4
        if stage == "solve":
            # the constrained acceleration is the minimiser of a strictly convex problem: both sides must agree
            d, o = get_dev(slab, "acc"), oracle.get("acc")
            scale = np.maximum(np.abs(o).max(axis=1, keepdims=True), 1.0)
            err = np.max(np.abs(d - o) / scale)
            assert err < 2e-3, f"acc after solve: {err}"
            _check_forces(slab, oracle, label)


@pytest.mark.parametrize("warm_steps,layout", [(0, "planes"), (20, "planes"), (20, "records")])
def test_franka_stage_parity(warm_steps, layout):
    rt = _rt()
    B = 128
    model, dm, slab = _franka_slab(B, warm_steps,
                                   layout=rt.LAYOUT_PLANES if layout == "planes" else rt.LAYOUT_ENV_RECORDS)
    _stage_parity(rt, model, slab, f"franka warm={warm_steps} {layout}")


def test_quadruped_stage_parity():
    rt = _rt()
    B = 64
    model = quadruped_model()
    dm = rt.DeviceModel(model)
    slab = rt.StateSlab(dm, B)
    slab["qpos"].copy_(torch.as_tensor(quadruped_initial_qpos(model, B).T, dtype=torch.float32))
    slab["ctrl_mode"].copy_(torch.as_tensor(quadruped_ctrl_mode(model, B).T, dtype=torch.int32))
    slab["ctrl_pos"].copy_(torch.as_tensor(quadruped_targets(model, B, 0).T, dtype=torch.float32))
    slab.forward_kinematics()
    slab.step(40, rt.RS_MODE_FUSED)  # feet land, legs move toward the targets
    torch.cuda.synchronize()
    _stage_parity(rt, model, slab, "quadruped")


def test_fused_step_forces_match_oracle():
    """One fused substep from identical state: constraint forces and contact forces against the FP32 restatement."""
    rt = _rt()
    for label, make in (("franka", lambda: _franka_slab(64, 20)[::2]),
                        ("box pile", lambda: _box_pile_slab(16, 60))):
        model, slab = make()
        oracle = Oracle(model, slab.n_envs, precision=32)
        slab_to_oracle(slab, oracle)
        slab.step(1, rt.RS_MODE_FUSED)
        torch.cuda.synchronize()
        oracle.step(1)
        print(f"[parity] fused step {label}", flush=True)
        assert np.array_equal(get_dev(slab, "n_contacts"), oracle.get("n_contacts"))
        assert np.array_equal(get_dev(slab, "n_con"), oracle.get("n_con"))
        _check_forces(slab, oracle, f"fused {label}")


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


def _box_pile_slab(B, steps):
    rt = _rt()
    model = box_pile_model()
    dm = rt.DeviceModel(model)
    slab = rt.StateSlab(dm, B)
    slab.step(steps, rt.RS_MODE_FUSED)  # boxes land and pile
    torch.cuda.synchronize()
    return model, slab


def test_box_pile_stage_parity():
    rt = _rt()
    model, slab = _box_pile_slab(32, 60)
    _stage_parity(rt, model, slab, "box pile")


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
