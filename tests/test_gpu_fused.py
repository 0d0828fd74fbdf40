"""The fused substep (one wavefront per environment) against the per-stage kernels and the FP32 CPU restatement.

Both device paths run the same arithmetic in the same order (no FMA contraction, sequential sums), so the fused step
must reproduce the per-stage step to the last bit on identical input; against the CPU restatement the tolerance is the
reference's single-precision one (5e-5 of the field's magnitude, tests/conftest.py:78 TOL_SINGLE).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from genesis_amd.scenes import box_pile_model, franka_initial_qpos, franka_model, franka_targets
from tests.oracle_lib import Oracle
from tests.parity import check_fused_matches_plain, get_dev, scaled_err, slab_to_oracle

pytestmark = pytest.mark.gpu

OUTPUTS = ["qpos", "vel", "acc", "qacc_ws", "links_pos", "links_quat", "cd_vel", "geoms_pos", "aabb_min",
           "mass_mat", "qf_constraint", "force", "n_contacts", "n_con", "n_broad", "errbits"]


def _rt():
    from genesis_amd import runtime

    runtime.load_library()
    return runtime


def _franka(B, warm=20, seed=0, layout=None):
    rt = _rt()
    model = franka_model()
    dm = rt.DeviceModel(model, rt.LAYOUT_ENV_RECORDS if layout is None else layout)
    slab = rt.StateSlab(dm, B)
    slab["qpos"].copy_(torch.as_tensor(franka_initial_qpos(model, B, seed).T, dtype=torch.float32))
    slab["ctrl_mode"].fill_(0)
    slab["ctrl_pos"].copy_(torch.as_tensor(franka_targets(model, B, 0).T, dtype=torch.float32))
    slab.forward_kinematics()
    if warm:
        slab.step(warm, rt.RS_MODE_PLAIN)
    torch.cuda.synchronize()
    return model, dm, slab


def _twin(rt, dm, slab):
    t = rt.StateSlab(dm, slab.n_envs)
    t.buffer.copy_(slab.buffer)
    return t


def _contacts(slab):
    n = get_dev(slab, "n_contacts")[:, 0].astype(int)
    out = []
    for name, w in (("contact_geom_a", 1), ("contact_geom_b", 1), ("contact_pos", 3), ("contact_normal", 3),
                    ("contact_penetration", 1), ("contact_force", 3)):
        v = get_dev(slab, name)
        out.append(np.concatenate([v[b, : n[b] * w] for b in range(len(n))]))
    return out


def _max_abs(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b))) if a.size else 0.0


@pytest.mark.parametrize("layout", ["records", "planes"])
@pytest.mark.parametrize("warm", [0, 20, 60])
def test_fused_step_equals_plain_step(warm, layout):
    rt = _rt()
    model, dm, a = _franka(256, warm, layout=rt.LAYOUT_PLANES if layout == "planes" else rt.LAYOUT_ENV_RECORDS)
    b = _twin(rt, dm, a)
    print(f"[fused] franka B=256 warm={warm} layout={layout} config={dm.fused_config}", flush=True)
    a.step(1, rt.RS_MODE_PLAIN)
    b.step(1, rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    for name in OUTPUTS:
        d = _max_abs(get_dev(a, name), get_dev(b, name))
        assert d == 0.0, f"{name}: fused differs from per-stage by {d:.3e}"
    for x, y in zip(_contacts(a), _contacts(b)):
        assert _max_abs(x, y) == 0.0


def test_state_layouts_agree():
    """The same 5 substeps on a slab with per-environment records and on one with planes: identical fields."""
    rt = _rt()
    _, _, a = _franka(96, 15, layout=rt.LAYOUT_ENV_RECORDS)
    _, _, b = _franka(96, 15, layout=rt.LAYOUT_PLANES)
    print("[fused] layouts: records (fused) vs planes (plain)", flush=True)
    a.step(5, rt.RS_MODE_FUSED)
    b.step(5, rt.RS_MODE_PLAIN)
    torch.cuda.synchronize()
    for name in OUTPUTS:
        d = _max_abs(get_dev(a, name), get_dev(b, name))
        assert d == 0.0, f"{name}: records vs planes differ by {d:.3e}"


def test_fused_launch_shapes_agree():
    rt = _rt()
    model, dm, ref = _franka(128, 30)
    want = _twin(rt, dm, ref)
    want.step(5, rt.RS_MODE_PLAIN)
    torch.cuda.synchronize()
    # every shape with the GJK/EPA polytope of single detections in the work field (the default), and the LDS
    # polytope (gjk_lds) on a few shapes
    shapes = [(epb, lds, sched, occ, False) for occ in (0, 3, 4) for sched in (True, False) for lds in (True, False)
              for epb in (1, 2, 4, 8)]
    shapes += [(1, True, True, 0, True), (2, True, False, 3, True), (4, False, True, 0, True)]
    for epb, lds, sched, occ, gjk_lds in shapes:
        try:
            dm.set_fused_config(epb, lds, sched, occ, gjk_lds)
        except rt.RigidStepError:
            continue
        print(f"[fused] launch shape epb={epb} lds={lds} schedule={sched} occupancy={occ} gjk_lds={gjk_lds}",
              flush=True)
        s = _twin(rt, dm, ref)
        s.step(5, rt.RS_MODE_FUSED)
        torch.cuda.synchronize()
        for name in ("qpos", "vel", "acc", "n_contacts"):
            assert _max_abs(get_dev(s, name), get_dev(want, name)) == 0.0, (epb, lds, sched, occ, gjk_lds, name)


@pytest.mark.parametrize("warm", [0, 40])
def test_fused_step_vs_fp32_oracle(warm):
    rt = _rt()
    B = 128
    model, dm, slab = _franka(B, warm)
    oracle = Oracle(model, B, precision=32)
    slab_to_oracle(slab, oracle)
    slab.step(1, rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    oracle.step(1)
    assert np.array_equal(get_dev(slab, "n_contacts"), oracle.get("n_contacts"))
    assert np.array_equal(get_dev(slab, "n_con"), oracle.get("n_con"))
    for name in ("qpos", "vel", "links_pos"):
        err = scaled_err(get_dev(slab, name), oracle.get(name))
        assert err < 5e-5, f"{name}: {err:.3e}"
    d, o = get_dev(slab, "acc"), oracle.get("acc")
    scale = np.maximum(np.abs(o).max(axis=1, keepdims=True), 1.0)
    assert np.max(np.abs(d - o) / scale) < 2e-3


def test_fused_box_pile_matches_plain():
    """120 dofs: the fused solver assembles its Hessian on the matrix cores (check_fused_matches_plain states the
    bar)."""
    rt = _rt()
    B = 32
    model = box_pile_model()
    dm = rt.DeviceModel(model)
    with pytest.raises(rt.RigidStepError):  # at most 4 environments (256 lanes) per workgroup above 24 dofs
        dm.set_fused_config(8, False, False, 0)
    a = rt.StateSlab(dm, B)
    a.step(60, rt.RS_MODE_PLAIN)
    torch.cuda.synchronize()
    b = _twin(rt, dm, a)
    print(f"[fused] box pile B={B} config={dm.fused_config}", flush=True)
    a.step(1, rt.RS_MODE_PLAIN)
    b.step(1, rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    check_fused_matches_plain(a, b, model)


def test_fused_quadruped_equals_plain_and_tracks_oracle():
    rt = _rt()
    from genesis_amd.scenes import quadruped_ctrl_mode, quadruped_initial_qpos, quadruped_model, quadruped_targets

    B = 128
    model = quadruped_model()
    dm = rt.DeviceModel(model)
    a = rt.StateSlab(dm, B)
    a["qpos"].copy_(torch.as_tensor(quadruped_initial_qpos(model, B).T, dtype=torch.float32))
    a["ctrl_mode"].copy_(torch.as_tensor(quadruped_ctrl_mode(model, B).T, dtype=torch.int32))
    a["ctrl_pos"].copy_(torch.as_tensor(quadruped_targets(model, B, 0).T, dtype=torch.float32))
    a.forward_kinematics()
    a.step(100, rt.RS_MODE_PLAIN)  # feet on the ground
    torch.cuda.synchronize()
    b = _twin(rt, dm, a)
    oracle = Oracle(model, B, precision=32)
    slab_to_oracle(a, oracle)
    print(f"[fused] quadruped B={B} config={dm.fused_config}", flush=True)
    a.step(1, rt.RS_MODE_PLAIN)
    b.step(1, rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    oracle.step(1)
    for name in OUTPUTS:
        d = _max_abs(get_dev(a, name), get_dev(b, name))
        assert d == 0.0, f"{name}: fused differs from per-stage by {d:.3e}"
    assert np.array_equal(get_dev(b, "n_contacts"), oracle.get("n_contacts"))
    for name in ("qpos", "vel"):
        err = scaled_err(get_dev(b, name), oracle.get(name))
        assert err < 5e-5, f"{name}: {err:.3e}"


@pytest.mark.parametrize("shape", [(2, True, False, 0), (1, False, False, 0), (2, True, False, 3), (4, False, True, 3)])
def test_fused_quadruped_launch_shapes_equal_plain(shape):
    """18 dofs: the second row-tile class (17 to 24 dofs), in each private-region placement and occupancy variant,
    over 5 substeps in contact (rank-1 updates and line searches on every row count the gait produces)."""
    rt = _rt()
    from genesis_amd.scenes import quadruped_ctrl_mode, quadruped_initial_qpos, quadruped_model, quadruped_targets

    B = 256
    model = quadruped_model()
    dm = rt.DeviceModel(model)
    dm.set_fused_config(*shape)
    a = rt.StateSlab(dm, B)
    a["qpos"].copy_(torch.as_tensor(quadruped_initial_qpos(model, B).T, dtype=torch.float32))
    a["ctrl_mode"].copy_(torch.as_tensor(quadruped_ctrl_mode(model, B).T, dtype=torch.int32))
    a["ctrl_pos"].copy_(torch.as_tensor(quadruped_targets(model, B, 0).T, dtype=torch.float32))
    a.forward_kinematics()
    a.step(80, rt.RS_MODE_PLAIN)
    torch.cuda.synchronize()
    b = _twin(rt, dm, a)
    print(f"[fused] quadruped B={B} launch shape {shape} config={dm.fused_config}", flush=True)
    a.step(5, rt.RS_MODE_PLAIN)
    b.step(5, rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    assert int(get_dev(a, "n_con").max()) > 0
    for name in OUTPUTS:
        d = _max_abs(get_dev(a, name), get_dev(b, name))
        assert d == 0.0, f"{shape} {name}: fused differs from per-stage by {d:.3e}"


def test_autotune_keeps_state_and_picks_a_shape():
    rt = _rt()
    model, dm, slab = _franka(512, 10)
    before = slab.buffer.clone()
    res = rt.autotune_fused(slab, steps=3)
    torch.cuda.synchronize()
    assert torch.equal(before, slab.buffer)
    assert res["envs_per_block"] in (1, 2, 4, 8)
    assert dm.fused_config["envs_per_block"] == res["envs_per_block"]
