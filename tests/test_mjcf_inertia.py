"""MJCF inertia from geoms (MuJoCo's inertiafromgeom) for a capsule given by `fromto`: mass, centre of mass and
principal moments against a direct numerical integration over the capsule's volume (uniform density)."""

from __future__ import annotations

import os
import tempfile

import numpy as np

from genesis_amd.model.mjcf import load_mjcf


def _capsule_by_sampling(a, b, r, density, n=1_500_000, seed=0):
    rng = np.random.default_rng(seed)
    lo = np.minimum(a, b) - r
    hi = np.maximum(a, b) + r
    p = rng.uniform(lo, hi, (n, 3))
    d = b - a
    t = np.clip((p - a) @ d / (d @ d), 0.0, 1.0)
    inside = np.linalg.norm(p - (a + t[:, None] * d), axis=1) <= r
    vol = np.prod(hi - lo) * inside.mean()
    q = p[inside]
    com = q.mean(axis=0)
    x = q - com
    inertia = density * vol * ((x * x).sum(1).mean() * np.eye(3) - (x.T @ x) / len(x))
    return density * vol, com, np.sort(np.linalg.eigvalsh(inertia))


def test_fromto_capsule_inertia():
    a, b, r, rho = np.array([0.0, 0.0, 0.0]), np.array([0.3, 0.0, -0.2]), 0.02, 500.0
    xml = (f'<mujoco><worldbody><body name="b" pos="0 0 1"><joint type="hinge" axis="0 1 0"/>'
           f'<geom type="capsule" fromto="{" ".join(map(str, a))} {" ".join(map(str, b))}" size="{r}" '
           f'density="{rho}"/></body></worldbody></mujoco>')
    with tempfile.NamedTemporaryFile("w", suffix=".xml", delete=False) as f:
        f.write(xml)
    try:
        link = load_mjcf(f.name, "t").links[0]
    finally:
        os.unlink(f.name)
    mass, com, moments = _capsule_by_sampling(a, b, r, rho)
    assert abs(link.inertial_mass - mass) < 5e-3 * mass
    np.testing.assert_allclose(link.inertial_pos, com, atol=2e-3)
    np.testing.assert_allclose(np.sort(np.diag(link.inertial_i)), moments, rtol=2e-2)
