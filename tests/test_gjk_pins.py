"""Pins of the GJK/EPA restatement (oracle/oracle_gjk.h, after collider/gjk.py:170-1915 and epa.py:969-1432).

The penetration depth of two convex sets is the distance from the origin to the boundary of their Minkowski
difference A - B, reached along that boundary's nearest facet normal. For polytopes that difference is the convex hull
of every vertex difference a_i - b_j, so scipy's qhull gives the exact depth and normal independently of GJK/EPA.
Smooth shapes are pinned by closed forms (sphere-sphere, crossed capsules). The contact normal must follow the MPR
path's convention, since the reference swaps one detector for the other on the same pair (narrowphase.py:726-851).
"""

from __future__ import annotations

import numpy as np
import pytest
from scipy.spatial import ConvexHull

from genesis_amd.model import transforms as tf
from genesis_amd.model.compiler import SolverOptions, compile_model
from genesis_amd.model.primitives import box_entity, capsule_entity, hull_entity, sphere_entity
from tests.oracle_lib import Oracle


def _oracle(ents, precision=64):
    m = compile_model(ents, SolverOptions())
    return m, Oracle(m, 1, precision=precision)


def _pose(p, q):
    q = np.asarray(q, np.float64)
    return np.concatenate([np.asarray(p, np.float64), q / np.linalg.norm(q)])


def _world(verts, pose):
    return verts @ tf.quat_to_R(pose[3:]).T + pose[:3]


def _minkowski_depth(va, vb):
    """(depth, unit normal n) of the nearest facet of hull(A - B) to the origin; n points out of A - B."""
    d = (va[:, None, :] - vb[None, :, :]).reshape(-1, 3)
    hull = ConvexHull(d)
    off = -hull.equations[:, 3]  # distance of each facet plane from the origin (positive: origin inside)
    k = int(np.argmin(off))
    return float(off[k]), hull.equations[k, :3], off


def _rand_quat(rng):
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


@pytest.mark.parametrize("precision", [64, 32])
def test_box_box_face_overlap_depth_and_normal(precision):
    # b rests on a, sunk by 0.03 m: the depth is 0.03 along z whichever detector runs
    m, o = _oracle([box_entity([0.4, 0.4, 0.4]), box_entity([0.3, 0.5, 0.2], pos=(0, 0, 1))], precision)
    pa = _pose([0, 0, 0], [1, 0, 0, 0])
    pb = _pose([0.02, -0.01, 0.2 + 0.1 - 0.03], [1, 0, 0, 0])
    col, n, p, pen = o.pair_contact(0, 1, pa, pb, "gjk")
    tol = 1e-9 if precision == 64 else 2e-6
    assert col and abs(pen - 0.03) < tol, pen
    assert abs(abs(n[2]) - 1.0) < tol
    # the contact point lies midway between the two faces, inside the overlap patch
    assert abs(p[2] - (0.2 - 0.015)) < 1e-6 and abs(p[0]) <= 0.17 and abs(p[1]) <= 0.2
    colm, nm, pm, penm = o.pair_contact(0, 1, pa, pb, "mpr")
    assert colm and abs(penm - 0.03) < 1e-5
    assert np.dot(n, nm) > 0.999, (n, nm)


def test_box_box_random_orientations_match_minkowski_hull():
    m, o = _oracle([box_entity([0.3, 0.2, 0.25]), box_entity([0.15, 0.35, 0.1], pos=(0, 0, 1))])
    va0, vb0 = m.arrays["geom_verts"][0], m.arrays["geom_verts"][1]
    rng = np.random.default_rng(7)
    n_checked = 0
    for _ in range(60):
        pa = _pose(rng.uniform(-0.05, 0.05, 3), _rand_quat(rng))
        pb = _pose(rng.uniform(-0.12, 0.12, 3), _rand_quat(rng))
        depth, nrm, off = _minkowski_depth(_world(va0, pa), _world(vb0, pb))
        col, n, p, pen = o.pair_contact(0, 1, pa, pb, "gjk")
        if depth <= 1e-6:
            assert not col or pen < 1e-6
            continue
        assert col, depth
        assert abs(pen - depth) < 1e-9 * max(1.0, depth) + 1e-12, (pen, depth)
        # the witness normal (from the point on a to the point on b) is minus the facet normal of A - B, unless two
        # facets tie for the minimum
        second = np.partition(off, 1)[1]
        if second - depth > 1e-6:
            assert np.dot(n, -nrm) > 1 - 1e-9, (n, nrm)
        n_checked += 1
    assert n_checked >= 30


def test_hull_hull_depth_matches_minkowski_hull_within_the_support_table():
    """Mesh supports come from the 180 x 180 direction table (support_field.py:139-181), an approximation of the
    exact argmax vertex, so the EPA depth of two hulls agrees with the exact one only up to the table's resolution."""
    rng = np.random.default_rng(3)
    worst = 0.0
    checked = 0
    for trial in range(8):
        ptsa = rng.normal(size=(14, 3)) * [0.12, 0.08, 0.1]
        ptsb = rng.normal(size=(10, 3)) * [0.07, 0.1, 0.06]
        m, o = _oracle([hull_entity(ptsa), hull_entity(ptsb, pos=(0, 0, 1))])
        va0, vb0 = m.arrays["geom_verts"][0], m.arrays["geom_verts"][1]
        for _ in range(6):
            pa = _pose(rng.uniform(-0.02, 0.02, 3), _rand_quat(rng))
            pb = _pose(rng.uniform(-0.08, 0.08, 3), _rand_quat(rng))
            depth, nrm, _ = _minkowski_depth(_world(va0, pa), _world(vb0, pb))
            col, n, p, pen = o.pair_contact(0, 1, pa, pb, "gjk")
            if depth < 5e-3:
                continue
            assert col
            worst = max(worst, abs(pen - depth) / depth)
            checked += 1
    assert checked >= 20
    assert worst < 0.05, worst


@pytest.mark.parametrize("precision", [64, 32])
def test_sphere_sphere_closed_form(precision):
    ra, rb = 0.1, 0.07
    m, o = _oracle([sphere_entity(ra), sphere_entity(rb, pos=(0, 0, 1))], precision)
    rng = np.random.default_rng(1)
    for _ in range(10):
        ca = rng.uniform(-0.1, 0.1, 3)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        d = rng.uniform(0.05, 0.16)
        cb = ca + d * u
        col, n, p, pen = o.pair_contact(0, 1, _pose(ca, [1, 0, 0, 0]), _pose(cb, [1, 0, 0, 0]), "gjk")
        assert col
        # EPA on smooth shapes stops within its tolerance of the curved boundary
        assert abs(pen - (ra + rb - d)) < 2e-3 * (ra + rb), (pen, ra + rb - d)
        assert abs(abs(np.dot(n, u)) - 1) < 2e-3
        mid = 0.5 * ((ca + ra * u) + (cb - rb * u))
        assert np.linalg.norm(p - mid) < 2e-3 * (ra + rb)
        colm, nm, pm, penm = o.pair_contact(0, 1, _pose(ca, [1, 0, 0, 0]), _pose(cb, [1, 0, 0, 0]), "mpr")
        assert np.dot(n, nm) > 0.99


def test_crossed_capsules_closed_form():
    r1, r2, l1, l2 = 0.05, 0.04, 0.4, 0.3
    m, o = _oracle([capsule_entity(r1, l1), capsule_entity(r2, l2, pos=(0, 0, 1))])
    qx = tf.quat_normalize(np.array([np.cos(np.pi / 4), 0, np.sin(np.pi / 4), 0]))  # axis along x
    qy = tf.quat_normalize(np.array([np.cos(np.pi / 4), np.sin(np.pi / 4), 0, 0]))  # axis along y
    for gap in (0.02, 0.05, 0.07):
        pa = _pose([0, 0, 0], qx)
        pb = _pose([0.03, -0.02, gap], qy)
        col, n, p, pen = o.pair_contact(0, 1, pa, pb, "gjk")
        assert col
        assert abs(pen - (r1 + r2 - gap)) < 2e-3 * (r1 + r2), (gap, pen)
        assert abs(abs(n[2]) - 1) < 1e-2


def test_separated_shapes_do_not_collide():
    m, o = _oracle([box_entity([0.2, 0.2, 0.2]), box_entity([0.2, 0.2, 0.2], pos=(0, 0, 1))])
    rng = np.random.default_rng(5)
    for _ in range(20):
        pa = _pose([0, 0, 0], _rand_quat(rng))
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        pb = _pose(0.4 * u, _rand_quat(rng))  # centres 0.4 apart > circumscribed radii 0.173 + 0.173
        col, n, p, pen = o.pair_contact(0, 1, pa, pb, "gjk")
        assert not col and pen == 0.0


def test_narrowphase_hands_deep_first_contacts_to_gjk():
    """A deep first contact (no cached normal) is detected again by GJK/EPA (narrowphase.py:726-734), so the
    narrowphase's first contact of the pair carries the exact EPA depth of the two boxes."""
    m, o = _oracle([box_entity([0.3, 0.3, 0.3]), box_entity([0.3, 0.3, 0.3], pos=(0.05, 0.02, 0.2),
                                                           quat=(0.96, 0.1, 0.2, 0.15))])
    o.forward_kinematics()
    for st in (4, 5, 6):  # AABB, broadphase, narrowphase
        o.stage(st)
    assert int(o.get("n_gjk_regime")[0, 0]) >= 1
    n = int(o.get("n_contacts")[0, 0])
    assert n >= 1
    qa = o.get("geoms_quat")[0].reshape(-1, 4)
    pa = o.get("geoms_pos")[0].reshape(-1, 3)
    va = _world(m.arrays["geom_verts"][0], np.concatenate([pa[0], qa[0]]))
    vb = _world(m.arrays["geom_verts"][1], np.concatenate([pa[1], qa[1]]))
    depth, _, _ = _minkowski_depth(va, vb)
    pen0 = o.get("contact_penetration")[0, 0]
    assert abs(pen0 - depth) < 1e-9, (pen0, depth)
