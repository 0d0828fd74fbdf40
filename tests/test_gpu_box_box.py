"""Box-box separating-axis detection on the device (genesis_amd/csrc/rs_collision.h box_box) against the CPU
restatement (oracle/oracle_collision.h box_box_contact, after collider/box_contact.py:100-785).

- 1024 random overlapping box pairs (face and edge contacts, one pair per environment): the narrowphase stage's
  contact lists equal the restatement's (counts, geoms exact; positions, normals, penetrations to the single-precision
  tolerance used by every stage-parity test);
- a box pile with box_box_detection=True: every stage against the restatement, and the fused step against the
  per-stage step (to the solver's tolerance: 120 dofs assemble the fused Hessian on the matrix cores);
- the reference's known answer, test_box_box_dynamics (tests/test_rigid_physics.py:1031-1053, run with
  box_box_detection=True by tests/utils.py:573), on the device: a 0.4 m cube dropped with a random tilt onto another
  settles at z = 0.6 +- 2e-3 with |v| < 1e-2.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from genesis_amd.model.compiler import SolverOptions, compile_model
from genesis_amd.model.primitives import box_entity, plane_entity
from genesis_amd.scenes import box_pile_model
from tests.oracle_lib import Oracle
from tests.parity import check_fused_matches_plain, get_dev, scaled_err, slab_to_oracle
from tests.test_gpu_parity import _masked, _stage_parity

pytestmark = pytest.mark.gpu


def _rt():
    from genesis_amd import runtime

    runtime.load_library()
    return runtime


def _twin(rt, dm, slab):
    t = rt.StateSlab(dm, slab.n_envs)
    t.buffer.copy_(slab.buffer)
    return t


def test_random_box_pairs_match_the_restatement():
    rt = _rt()
    m = compile_model([box_entity([0.3, 0.2, 0.25], name="a"), box_entity([0.15, 0.35, 0.1], pos=(0, 0, 1), name="b")],
                      SolverOptions(gravity=(0.0, 0.0, 0.0), box_box_detection=True))
    assert m.n_contacts_per_pair == 16
    B = 1024
    rng = np.random.default_rng(3)
    q = np.zeros((B, 14))
    for b in range(B):
        qa, qb = rng.normal(size=4), rng.normal(size=4)
        q[b, :3] = rng.uniform(-0.05, 0.05, 3)
        q[b, 3:7] = qa / np.linalg.norm(qa)
        q[b, 7:10] = rng.uniform(-0.2, 0.2, 3)
        q[b, 10:] = qb / np.linalg.norm(qb)
    # a quarter of the pairs axis-aligned and stacked (face contacts with 4-8 points)
    for b in range(0, B, 4):
        q[b, :7] = [0, 0, 0, 1, 0, 0, 0]
        yaw = rng.uniform(0, np.pi)
        q[b, 10:] = [np.cos(yaw / 2), 0, 0, np.sin(yaw / 2)]
        q[b, 7:10] = [*rng.uniform(-0.1, 0.1, 2), 0.125 + 0.05 - rng.uniform(0.001, 0.02)]
    dm = rt.DeviceModel(m)
    slab = rt.StateSlab(dm, B)
    slab["qpos"].copy_(torch.as_tensor(q.T, dtype=torch.float32))
    slab.forward_kinematics()
    torch.cuda.synchronize()
    oracle = Oracle(m, B, precision=32)
    for stage in ("aabb", "broadphase", "narrowphase"):
        slab_to_oracle(slab, oracle)
        slab.stage(stage)
        oracle.stage(rt.STAGE_ID[stage])
    torch.cuda.synchronize()
    nc_d, nc_o = get_dev(slab, "n_contacts")[:, 0], oracle.get("n_contacts")[:, 0]
    print(f"[box-box] {int((nc_o > 0).sum())} of {B} pairs in contact, {int(nc_o.sum())} points, "
          f"most {int(nc_o.max())}", flush=True)
    assert (nc_o > 0).sum() > B // 3 and nc_o.max() >= 8
    assert np.array_equal(nc_d, nc_o), f"contact counts differ at {np.nonzero(nc_d != nc_o)[0][:10]}"
    for f in ("contact_geom_a", "contact_geom_b", "contact_pair"):
        d, o = _masked(slab, oracle, f, "n_contacts", 1)
        assert np.array_equal(d, o), f
    for f in ("contact_pos", "contact_normal"):
        d, o = _masked(slab, oracle, f, "n_contacts", 3)
        err = scaled_err(d, o)
        print(f"[box-box] {f}: scaled error {err:.3e}", flush=True)
        assert err < 1e-4, f
    d, o = _masked(slab, oracle, "contact_penetration", "n_contacts", 1)
    assert np.max(np.abs(d - o)) < 5e-5


def _pile(B, steps, mode):
    rt = _rt()
    model = box_pile_model(box_box_detection=True)
    assert model.n_contacts_per_pair == 16
    dm = rt.DeviceModel(model)
    slab = rt.StateSlab(dm, B)
    slab.step(steps, rt.RS_MODE_FUSED if mode == "fused" else rt.RS_MODE_PLAIN)
    torch.cuda.synchronize()
    return rt, model, dm, slab


def test_box_pile_stage_parity_with_box_box_detection():
    rt, model, _, slab = _pile(32, 60, "fused")
    _stage_parity(rt, model, slab, "box pile (box-box)")


def test_fused_box_pile_with_box_box_detection_matches_plain():
    rt, model, dm, a = _pile(32, 60, "plain")
    b = _twin(rt, dm, a)
    print(f"[box-box] box pile fused config={dm.fused_config}", flush=True)
    a.step(1, rt.RS_MODE_PLAIN)
    b.step(1, rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    check_fused_matches_plain(a, b, model)


BOX_BOX = """<mujoco model="two_boxes"><option timestep="0.01"/>
<default><geom contype="1" conaffinity="1" condim="3" friction="1. 0.5 0.5"/></default>
<worldbody>
  <body name="box1" pos="0 0 0.2"><geom type="box" size="0.2 0.2 0.2"/><joint name="root1" type="free"/></body>
  <body name="box2" pos="0 0 0.8"><geom type="box" size="0.2 0.2 0.2"/><joint name="root2" type="free"/></body>
</worldbody></mujoco>"""


def _quat_mul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


@pytest.mark.parametrize("solver", [0, 1])
def test_reference_box_box_stack_settles_on_the_device(tmp_path, solver):
    import os

    from genesis_amd.model.mjcf import load_mjcf

    rt = _rt()
    path = os.path.join(str(tmp_path), "bb.xml")
    with open(path, "w") as f:
        f.write(BOX_BOX)
    m = compile_model([plane_entity(friction=1.0), load_mjcf(path, "boxes")],
                      SolverOptions(dt=0.01, box_box_detection=True, constraint_solver=solver))
    B = 256
    rng = np.random.default_rng(6)
    q = np.zeros((B, 14))
    for b in range(B):
        e = np.array([*(0.15 * rng.random(2)), np.pi * rng.random()])
        cz, sz = np.cos(e / 2), np.sin(e / 2)
        qx, qy, qz = np.array([cz[0], sz[0], 0, 0]), np.array([cz[1], 0, sz[1], 0]), np.array([cz[2], 0, 0, sz[2]])
        q[b] = [0, 0, 0.2, 1, 0, 0, 0, 0, 0, 0.65 + 0.1 * rng.random(), *_quat_mul(_quat_mul(qz, qy), qx)]
    dm = rt.DeviceModel(m)
    slab = rt.StateSlab(dm, B)
    slab["qpos"].copy_(torch.as_tensor(q.T, dtype=torch.float32))
    slab.forward_kinematics()
    slab.step(101, rt.RS_MODE_FUSED)
    for _ in range(9):
        slab.step(1, rt.RS_MODE_FUSED)
        torch.cuda.synchronize()
        assert np.abs(get_dev(slab, "vel")).max() < 1e-2
    z = get_dev(slab, "qpos")[:, 9]
    print(f"[box-box] stack solver={solver}: top cube z in [{z.min():.5f}, {z.max():.5f}]", flush=True)
    assert np.allclose(z, 0.6, atol=2e-3)
    assert int(get_dev(slab, "errbits").max()) == 0
