"""Device GJK/EPA (genesis_amd/csrc/rs_gjk.h) against the CPU restatement (oracle/oracle_gjk.h), one detection per item.

Both device forms are checked: the fused kernel's (a whole wavefront per detection, polytope in LDS, run again in
memory when it outgrows that) and the per-stage kernels' (one lane per detection, polytope in memory). Items:

- every hull/box pair of the Franka model at states reached by stepping random joint configurations toward random
  targets (most pairs apart: GJK alone; the overlapping ones also run EPA), and the overlapping ones pushed 5, 10 and
  20 mm deeper along their contact normal (the deep overlaps the MPR path hands over);
- smooth shapes (spheres, capsules) in deep overlap, whose EPA runs long enough to outgrow the LDS polytope.

Contact flags must be equal; normal, position and penetration agree to 1e-5 (the support-table lookup goes through
atan2/acos, whose last bit may differ between the device and the host libm).
The narrowphase's use of the hand-over (narrowphase.py:726-851) is covered by the stage-parity tests on Franka states.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from genesis_amd.model.compiler import SolverOptions, compile_model
from genesis_amd.model.primitives import capsule_entity, sphere_entity
from genesis_amd.scenes import franka_initial_qpos, franka_model, franka_targets
from tests.oracle_lib import Oracle

pytestmark = pytest.mark.gpu

GT_BOX, GT_MESH = 5, 6


def _franka_cases():
    m = franka_model()
    vp = np.asarray(m.arrays["valid_pairs"]).reshape(-1, 2)
    gt = np.asarray(m.arrays["geom_type"]).ravel()
    B = 16
    o = Oracle(m, B, precision=32)
    o.set("qpos", franka_initial_qpos(m, B, 0))
    o.set("ctrl_mode", np.zeros((B, m.n_dofs)))
    o.forward_kinematics()
    o.set("ctrl_pos", franka_targets(m, B, 0))
    cases = []
    for _ in range(6):
        o.step(10, 8)
        gp = o.get("geoms_pos").reshape(B, -1, 3)
        gq = o.get("geoms_quat").reshape(B, -1, 4)
        for b in range(B):
            for ga, gb in vp:
                if gt[ga] in (GT_BOX, GT_MESH) and gt[gb] in (GT_BOX, GT_MESH):
                    cases.append((int(ga), int(gb), np.concatenate([gp[b, ga], gq[b, ga]]),
                                  np.concatenate([gp[b, gb], gq[b, gb]])))
    deep = []
    for ga, gb, pa, pb in cases:
        col, n, _, _ = o.pair_contact(ga, gb, pa, pb, "gjk")
        if col:
            for d in (0.005, 0.01, 0.02):
                q = pb.copy()
                q[:3] += d * n
                deep.append((ga, gb, pa, q))
    rng = np.random.default_rng(0)
    keep = rng.choice(len(cases), size=min(1500, len(cases)), replace=False)
    return m, o, [cases[i] for i in keep] + deep


def _smooth_cases():
    ents = [sphere_entity(0.1), sphere_entity(0.07, pos=(0, 0, 1)), capsule_entity(0.05, 0.3, pos=(1, 0, 0)),
            capsule_entity(0.04, 0.2, pos=(2, 0, 0))]
    m = compile_model(ents, SolverOptions())
    o = Oracle(m, 1, precision=32)
    rng = np.random.default_rng(1)
    cases = []
    for _ in range(40):
        for ga, gb in ((0, 1), (0, 2), (2, 3), (1, 3)):
            q1, q2 = rng.normal(size=4), rng.normal(size=4)
            pa = np.concatenate([rng.uniform(-0.02, 0.02, 3), q1 / np.linalg.norm(q1)])
            pb = np.concatenate([rng.uniform(-0.06, 0.06, 3), q2 / np.linalg.norm(q2)])
            cases.append((ga, gb, pa, pb))
    return m, o, cases


def _compare(m, o, cases, label):
    from genesis_amd import runtime as rt

    rt.load_library()
    dm = rt.DeviceModel(m)
    ga = np.array([c[0] for c in cases], np.int32)
    gb = np.array([c[1] for c in cases], np.int32)
    pa = np.stack([c[2] for c in cases]).astype(np.float32)
    pb = np.stack([c[3] for c in cases]).astype(np.float32)
    ref = [o.pair_contact(int(a), int(b), x.astype(np.float64), y.astype(np.float64), "gjk")
           for a, b, x, y in zip(ga, gb, pa, pb)]
    rhit = np.array([r[0] for r in ref], np.int32)
    rout = np.array([np.concatenate([r[1], r[2], [r[3]]]) for r in ref])
    print(f"[gjk] {label}: {len(cases)} items, {int(rhit.sum())} in contact", flush=True)
    for lanes in (64, 1):
        hit, out = rt.pair_contact(dm, ga, gb, pa, pb, rt.RS_PAIR_GJK, lanes)
        torch.cuda.synchronize()
        hit, out = hit.cpu().numpy(), out.cpu().numpy()
        assert np.array_equal(hit, rhit), f"{label} lanes={lanes}: contact flags differ at {np.nonzero(hit != rhit)[0][:10]}"
        both = rhit == 1
        err = float(np.abs(out[both] - rout[both]).max()) if both.any() else 0.0
        print(f"[gjk] {label} lanes={lanes}: max |device - restatement| {err:.3e}", flush=True)
        assert err <= 1e-5, f"{label} lanes={lanes}: {err}"
    return int(rhit.sum())


def test_franka_hull_and_box_pairs_match_the_restatement():
    m, o, cases = _franka_cases()
    n_hit = _compare(m, o, cases, "franka")
    assert n_hit >= 100


def test_deep_smooth_overlaps_match_the_restatement():
    m, o, cases = _smooth_cases()
    n_hit = _compare(m, o, cases, "spheres and capsules")
    assert n_hit >= 50


def test_mpr_detections_match_the_restatement():
    """The same items through MPR without a warm-start direction (mpr.py:769-818), one lane per item."""
    from genesis_amd import runtime as rt

    m, o, cases = _franka_cases()
    dm = rt.DeviceModel(m)
    ga = np.array([c[0] for c in cases], np.int32)
    gb = np.array([c[1] for c in cases], np.int32)
    pa = np.stack([c[2] for c in cases]).astype(np.float32)
    pb = np.stack([c[3] for c in cases]).astype(np.float32)
    ref = [o.pair_contact(int(a), int(b), x.astype(np.float64), y.astype(np.float64), "mpr")
           for a, b, x, y in zip(ga, gb, pa, pb)]
    rhit = np.array([r[0] for r in ref], np.int32)
    rout = np.array([np.concatenate([r[1], r[2], [r[3]]]) for r in ref])
    hit, out = rt.pair_contact(dm, ga, gb, pa, pb, rt.RS_PAIR_MPR, 1)
    torch.cuda.synchronize()
    hit, out = hit.cpu().numpy(), out.cpu().numpy()
    assert np.array_equal(hit, rhit)
    both = rhit == 1
    err = float(np.abs(out[both] - rout[both]).max())
    print(f"[gjk] mpr: {int(both.sum())} contacts, max |device - restatement| {err:.3e}", flush=True)
    assert err <= 1e-5


def test_invalid_items_are_flagged():
    from genesis_amd import runtime as rt

    m = franka_model()
    dm = rt.DeviceModel(m)
    pose = np.array([[0, 0, 0, 1, 0, 0, 0]] * 3, np.float32)
    # geom 0 is the plane; index m.n_geoms is out of range
    hit, _ = rt.pair_contact(dm, [0, 1, m.n_geoms], [1, 2, 1], pose, pose, rt.RS_PAIR_GJK, 64)
    assert hit.cpu().tolist()[0] == -1 and hit.cpu().tolist()[2] == -1
    assert hit.cpu().tolist()[1] in (0, 1)
