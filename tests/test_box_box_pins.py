"""Pins of the box-box separating-axis detection (oracle/oracle_collision.h box_box_contact, after
collider/box_contact.py:100-785, MuJoCo's engine_collision_box.c as the reference restates it).

- Closed forms: a face contact (a small box sunk into a larger one's top face) and an edge-edge crossing.
- Random overlaps: the contact normal is the facet normal of hull(A - B) nearest the origin (the exact separating
  axis of least overlap, from scipy's qhull), with the MPR path's sign; every point lies in the overlap region; no
  point is deeper than the exact depth.
- The reference's own known answer: test_box_box_dynamics (tests/test_rigid_physics.py:1031-1053), which runs with
  box_box_detection=True (tests/utils.py:573).

The reference reports, for a face contact, the height of the clipped point above the face after halving it
(box_contact.py:417-434: depth[n] keeps the full height, box_points[n][2] is halved and is what becomes the
distance), so a face contact of depth d reports penetration d / 2 at the midpoint; edge contacts report the full
depth (box_contact.py:740). The pins below state that.
"""

from __future__ import annotations

import numpy as np
import pytest
from scipy.spatial import ConvexHull

from genesis_amd.model import transforms as tf
from genesis_amd.model.compiler import SolverOptions, compile_model
from genesis_amd.model.primitives import box_entity, plane_entity
from genesis_amd.runtime import STAGE_ID
from tests.oracle_lib import Oracle


def _pair_model(sa, sb, box_box=True):
    return compile_model([box_entity(sa, name="a"), box_entity(sb, pos=(0, 0, 1), name="b")],
                         SolverOptions(gravity=(0.0, 0.0, 0.0), box_box_detection=box_box))


def _detect(m, poses, precision=64, box_box=True):
    """Contacts (pos, normal, penetration) of each pose pair [(pa, pb), ...], one env per pair."""
    o = Oracle(m, len(poses), precision=precision)
    q = np.zeros((len(poses), 14))
    for b, (pa, pb) in enumerate(poses):
        q[b, :7], q[b, 7:] = pa, pb
    o.set("qpos", q)
    o.forward_kinematics()
    for s in ("aabb", "broadphase", "narrowphase"):
        o.stage(STAGE_ID[s])
    out = []
    for b in range(len(poses)):
        n = int(o.get("n_contacts")[b, 0])
        out.append((o.get("contact_pos")[b, : 3 * n].reshape(n, 3), o.get("contact_normal")[b, : 3 * n].reshape(n, 3),
                    o.get("contact_penetration")[b, :n]))
    return out


def _pose(p, q=(1, 0, 0, 0)):
    q = np.asarray(q, np.float64)
    return np.concatenate([np.asarray(p, np.float64), q / np.linalg.norm(q)])


def _axis_quat(axis, angle):
    axis = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    return np.concatenate([[np.cos(angle / 2)], np.sin(angle / 2) * axis])


def test_capacity_is_sixteen_points_per_pair():
    """collider.py:145-150: box-box clipping raises the per-pair capacity to 16 when the scene has two boxes."""
    assert _pair_model([0.2] * 3, [0.2] * 3).n_contacts_per_pair == 16
    assert _pair_model([0.2] * 3, [0.2] * 3, box_box=False).n_contacts_per_pair == 5


@pytest.mark.parametrize("precision", [64, 32])
def test_face_contact_closed_form(precision):
    """b (0.3 x 0.3 x 0.2) sunk d into the top face of a (0.4 x 0.4 x 0.2): four contacts at b's bottom corners,
    normal -z (the MPR path's sign for a below b), penetration d / 2 at z = top - d / 2."""
    d = 0.01
    m = _pair_model([0.4, 0.4, 0.2], [0.3, 0.3, 0.2])
    cx, cy = 0.02, -0.01
    (pos, nrm, pen), = _detect(m, [(_pose([0, 0, 1.0]), _pose([cx, cy, 1.2 - d]))], precision)
    tol = 1e-9 if precision == 64 else 2e-6
    assert len(pen) == 4
    want = {(cx + sx * 0.15, cy + sy * 0.15) for sx in (-1, 1) for sy in (-1, 1)}
    got = {(round(x, 6), round(y, 6)) for x, y in pos[:, :2]}
    assert got == {(round(x, 6), round(y, 6)) for x, y in want}
    assert np.allclose(pos[:, 2], 1.1 - d / 2, atol=tol)
    assert np.allclose(pen, d / 2, atol=tol)
    assert np.allclose(nrm, [0, 0, -1], atol=tol)
    # the MPR path's normal for the same pair has the same sign
    (_, nrm_mpr, pen_mpr), = _detect(_pair_model([0.4, 0.4, 0.2], [0.3, 0.3, 0.2], box_box=False),
                                     [(_pose([0, 0, 1.0]), _pose([cx, cy, 1.2 - d]))], precision)
    assert np.dot(nrm[0], nrm_mpr[0]) > 0.999


def test_face_contact_overhang_clips_to_the_lower_face():
    """b overhangs a's top face on +x: the contact patch is the overlap rectangle, its two outer corners on a's edge."""
    d = 0.004
    m = _pair_model([0.4, 0.4, 0.2], [0.3, 0.3, 0.2])
    (pos, nrm, pen), = _detect(m, [(_pose([0, 0, 1.0]), _pose([0.15, 0.0, 1.2 - d]))])
    xy = np.round(pos[:, :2], 9)
    assert len(pen) == 4
    assert {tuple(p) for p in xy} == {(0.0, -0.15), (0.0, 0.15), (0.2, -0.15), (0.2, 0.15)}
    assert np.allclose(pen, d / 2, atol=1e-12)
    assert np.allclose(nrm, [0, 0, -1], atol=1e-12)


@pytest.mark.parametrize("precision", [64, 32])
def test_edge_edge_crossing_closed_form(precision):
    """a turned 45 deg about x (top edge along x) and b turned 45 deg about y (bottom edge along y) above it, the edges
    crossing at the z axis and overlapping by d: the contacts sit at the crossing, the deepest at depth d, midway
    between the edges, normal -z."""
    h, d = 0.1, 0.006
    m = _pair_model([2 * h] * 3, [2 * h] * 3)
    top = h * np.sqrt(2)
    pa = _pose([0, 0, 0], _axis_quat([1, 0, 0], np.pi / 4))
    pb = _pose([0, 0, 2 * top - d], _axis_quat([0, 1, 0], np.pi / 4))
    (pos, nrm, pen), = _detect(m, [(pa, pb)], precision)
    tol = 1e-9 if precision == 64 else 5e-6
    assert len(pen) >= 1
    k = int(np.argmax(pen))
    assert abs(pen[k] - d) < tol, pen
    assert np.allclose(pos[k], [0, 0, top - d / 2], atol=tol), pos
    assert np.allclose(nrm, [0, 0, -1], atol=tol)
    assert np.all(pen <= d + tol)
    assert np.all(np.abs(pos[:, :2]) < d + tol)


def _minkowski(va, vb):
    dif = (va[:, None, :] - vb[None, :, :]).reshape(-1, 3)
    hull = ConvexHull(dif)
    off = -hull.equations[:, 3]
    k = int(np.argmin(off))
    return float(off[k]), hull.equations[k, :3], off


def _inside(p, pose, half, tol):
    local = (p - pose[:3]) @ tf.quat_to_R(pose[3:])
    return np.all(np.abs(local) <= half + tol)


def test_random_overlaps_take_the_least_overlap_axis():
    sa, sb = np.array([0.3, 0.2, 0.25]), np.array([0.15, 0.35, 0.1])
    m = _pair_model(sa, sb)
    va0, vb0 = m.arrays["geom_verts"][0], m.arrays["geom_verts"][1]
    rng = np.random.default_rng(11)
    poses, info = [], []
    while len(poses) < 120:
        qa, qb = rng.normal(size=4), rng.normal(size=4)
        pa = _pose(rng.uniform(-0.05, 0.05, 3), qa)
        pb = _pose(rng.uniform(-0.2, 0.2, 3), qb)
        depth, nrm, off = _minkowski(va0 @ tf.quat_to_R(pa[3:]).T + pa[:3], vb0 @ tf.quat_to_R(pb[3:]).T + pb[:3])
        poses.append((pa, pb))
        info.append((depth, nrm, np.partition(off, 1)[1] - depth))
    res = _detect(m, poses)
    mpr = _detect(_pair_model(sa, sb, box_box=False), poses)
    n_hit = 0
    for (pa, pb), (depth, nrm, gap), (pos, n, pen), (_, n_m, _) in zip(poses, info, res, mpr):
        if depth < -1e-9:
            assert len(pen) == 0
            continue
        if depth < 1e-6:
            continue
        assert len(pen) >= 1, depth
        n_hit += 1
        # one normal per pair, -(facet normal of A - B) unless two facets tie
        assert np.allclose(n, n[0], atol=1e-12)
        if gap > 1e-6:
            assert np.dot(n[0], nrm) > 1 - 1e-9, (n[0], nrm)
        if len(n_m):
            assert np.dot(n[0], n_m[0]) > 0.9
        assert np.all(pen <= depth + 1e-9) and np.all(pen >= -1e-9)
        for p in pos:
            assert _inside(p, pa, sa / 2, depth + 1e-9) and _inside(p, pb, sb / 2, depth + 1e-9)
    assert n_hit >= 60


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


@pytest.mark.parametrize("solver", ["newton", "cg"])
@pytest.mark.parametrize("precision", [64, 32])
def test_reference_box_box_stack_settles_with_box_box_detection(tmp_path, precision, solver):
    """test_box_box_dynamics (tests/test_rigid_physics.py:1031-1053, box_box_detection=True): a 0.4 m cube dropped
    with a random tilt onto another settles with its centre at z = 0.6 +- 2e-3 and |v| < 1e-2 after 100 steps."""
    import os

    from genesis_amd.model.mjcf import load_mjcf

    path = os.path.join(str(tmp_path), "bb.xml")
    with open(path, "w") as f:
        f.write(BOX_BOX)
    ent = load_mjcf(path, "boxes")
    m = compile_model([plane_entity(friction=1.0), ent],
                      SolverOptions(dt=0.01, box_box_detection=True, constraint_solver=1 if solver == "newton" else 0))
    assert m.n_contacts_per_pair == 16
    n = 6
    o = Oracle(m, n, precision=precision)
    rng = np.random.default_rng(5)
    q = np.zeros((n, 14))
    for b in range(n):
        e = np.array([*(0.15 * rng.random(2)), np.pi * rng.random()])
        cz, sz = np.cos(e / 2), np.sin(e / 2)
        qx, qy, qz = np.array([cz[0], sz[0], 0, 0]), np.array([cz[1], 0, sz[1], 0]), np.array([cz[2], 0, 0, sz[2]])
        q[b] = [0, 0, 0.2, 1, 0, 0, 0, 0, 0, 0.65 + 0.1 * rng.random(), *_quat_mul(_quat_mul(qz, qy), qx)]
    o.set("qpos", q)
    o.forward_kinematics()
    o.step(101, 1)
    for _ in range(9):
        o.step(1, 1)
        assert np.abs(o.get("vel")).max() < 1e-2
    assert np.allclose(o.get("qpos")[:, 9], 0.6, atol=2e-3)
    # the resting contact between the cubes is a face contact: the vertices of the overlap of two equal squares turned
    # against each other (4 to 8 points)
    for s in ("aabb", "broadphase", "narrowphase"):
        o.stage(STAGE_ID[s])
    ga = o.get("contact_geom_a")
    gb = o.get("contact_geom_b")
    for b in range(n):
        k = int(o.get("n_contacts")[b, 0])
        between = np.sum((ga[b, :k] == 1) & (gb[b, :k] == 2))
        assert 4 <= between <= 8, between


def test_reference_many_boxes_stay_stacked():
    """test_many_boxes_dynamics (tests/test_rigid_physics.py:1056-1111, box_box_detection=True, no initial velocity):
    125 unit cubes in a 5 x 5 x 5 grid (spacing 1 - 1e-3, bottom layer centred at z = 0.5) on a plane, 1000 collision
    pairs, 150 steps of 0.01 s: over the last 50 steps the mean |qvel| of a cube stays below 0.05, and at the end every
    cube is within 0.05 of its start position and 0.03 of its start orientation."""
    ents = [plane_entity()]
    start = []
    for n in range(125):
        i, j, k = n // 25, (n // 5) % 5, n % 5
        start.append((i * (1.0 - 1e-3), j * (1.0 - 1e-3), k * (1.0 - 1e-3) + 0.5))
        ents.append(box_entity([1.0, 1.0, 1.0], pos=start[-1], name=f"box{n}"))
    m = compile_model(ents, SolverOptions(dt=0.01, max_collision_pairs=1000, box_box_detection=True))
    o = Oracle(m, 1, precision=64)
    for step in range(150):
        o.step(1, 1)
        if step > 100:
            v = o.get("vel").reshape(125, 6)
            assert np.linalg.norm(v, axis=1).mean() < 0.05, step
    q = o.get("qpos").reshape(125, 7)
    assert np.abs(q[:, :3] - np.array(start)).max() < 0.05
    # orientation: the reference compares the rotation-vector part of the dofs with 0 (atol 0.03)
    w = np.clip(np.abs(q[:, 3]), 0.0, 1.0)
    assert (2 * np.arccos(w)).max() < 0.03
