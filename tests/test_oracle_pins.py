"""Pins of the CPU restatement (oracle/) against what mechanics and geometry fix independently of its code.

The reference cannot run here and ships no golden vectors for this path, so each oracle stage is checked against a
closed form, an identity that a plausible mistake breaks (a dropped term, a wrong sign or frame, a transposed
operand), brute force, or one of the reference's own known-answer tests. All run on the CPU at FP64 unless stated.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

from genesis_amd.model.compiler import SolverOptions, compile_model
from genesis_amd.model.mjcf import load_mjcf
from genesis_amd.model.primitives import box_entity, capsule_entity, plane_entity, sphere_entity
from genesis_amd.runtime import STAGE_ID
from genesis_amd.scenes import box_pile_model, franka_initial_qpos, franka_model
from tests.oracle_lib import Oracle

G = 9.81


def _quat_mul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def _quat_R(q):
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _rotvec_quat(rv):
    th = np.linalg.norm(rv)
    if th < 1e-300:
        return np.array([1.0, 0, 0, 0])
    return np.concatenate([[np.cos(th / 2)], np.sin(th / 2) * rv / th])


def _write(tmp_path, name, text):
    p = os.path.join(str(tmp_path), name)
    with open(p, "w") as f:
        f.write(text)
    return p


PENDULUM = """<mujoco model="pendulum"><option timestep="0.01"/><worldbody>
  <body name="arm" pos="0 0 1"><joint name="hinge" type="hinge" axis="0 1 0"/>
    <inertial pos="0 0 -0.5" mass="2" diaginertia="0.01 0.02 0.03"/>
    <geom type="sphere" size="0.05" pos="0 0 -0.5" contype="0" conaffinity="0"/></body>
</worldbody></mujoco>"""


# ------------------------------------------------------------------------------------------- integration
def test_free_fall_is_semi_implicit_euler():
    """v_n = -g n dt and z_n = z0 - g dt^2 n (n + 1) / 2 (semi-implicit: position uses the updated velocity)."""
    m = compile_model([sphere_entity(0.05, pos=(0, 0, 5.0))], SolverOptions(dt=0.01))
    o = Oracle(m, 2, precision=64)
    n = 40
    o.step(n, 1)
    q, v = o.get("qpos"), o.get("vel")
    dt = 0.01
    assert np.allclose(v[:, 2], -G * n * dt, rtol=0, atol=1e-12)
    assert np.allclose(q[:, 2], 5.0 - G * dt * dt * n * (n + 1) / 2, rtol=0, atol=1e-12)
    assert np.allclose(q[:, :2], 0.0, atol=1e-15) and np.allclose(q[:, 3:], [1, 0, 0, 0], atol=1e-15)


def test_torque_free_sphere_rotates_in_its_body_frame():
    """A sphere has an isotropic inertia, so its body angular velocity w stays constant and the orientation after n
    steps is q0 * exp(n dt w) (body-frame increment, right multiplication)."""
    q0 = np.array([0.8, 0.2, -0.4, 0.4])
    q0 /= np.linalg.norm(q0)
    m = compile_model([sphere_entity(0.1, pos=(0, 0, 0), quat=q0)], SolverOptions(dt=0.01, gravity=(0.0, 0.0, 0.0)))
    o = Oracle(m, 1, precision=64)
    w = np.array([0.7, -1.3, 2.1])
    vel = np.zeros((1, 6))
    vel[0, 3:] = w
    o.set("vel", vel)
    n = 25
    o.step(n, 1)
    want = _quat_mul(q0, _rotvec_quat(w * 0.01 * n))
    got = o.get("qpos")[0, 3:]
    assert np.allclose(got * np.sign(got[0]), want * np.sign(want[0]), atol=1e-9)
    assert np.allclose(o.get("vel")[0, 3:], w, atol=1e-12)


def test_pendulum_acceleration_closed_form(tmp_path):
    """theta'' = -m g l sin(theta) / (I_yy + m l^2) for a hinge about y with the COM at distance l below it."""
    m = compile_model([load_mjcf(_write(tmp_path, "p.xml", PENDULUM), "pend")], SolverOptions(dt=0.01))
    o = Oracle(m, 4, precision=64)
    th = np.array([[0.3], [-1.0], [2.0], [3.0]])
    o.set("qpos", th)
    o.forward_kinematics()
    o.step(1, 1)
    want = -2.0 * G * 0.5 * np.sin(th) / (0.02 + 2.0 * 0.25)
    assert np.allclose(o.get("acc"), want, rtol=1e-13, atol=0)


# ------------------------------------------------------------------------------------- kinematics / CRBA
def _franka_state(n, seed):
    m = franka_model()
    o = Oracle(m, n, precision=64)
    rng = np.random.default_rng(seed)
    q = franka_initial_qpos(m, n, seed)
    v = rng.normal(size=(n, m.n_dofs))
    o.set("qpos", q)
    o.set("vel", v)
    o.set("ctrl_mode", np.full((n, m.n_dofs), 2.0))  # FORCE: no velocity gain folded into M
    o.forward_kinematics()
    return m, o, q, v


def _link_frames(m, o):
    """Per env: link origins, link rotations, COM positions and world inertias (from the oracle's FK)."""
    A = m.arrays
    pos = o.get("links_pos").reshape(o.n_envs, -1, 3)
    quat = o.get("links_quat").reshape(o.n_envs, -1, 4)
    out = []
    for b in range(o.n_envs):
        Rs = [_quat_R(quat[b, l]) for l in range(m.n_links)]
        com = np.array([pos[b, l] + Rs[l] @ A["link_inertial_pos"][l] for l in range(m.n_links)])
        Iw = []
        for l in range(m.n_links):
            Ri = Rs[l] @ _quat_R(A["link_inertial_quat"][l])
            Iw.append(Ri @ A["link_inertial_i"][l] @ Ri.T)
        out.append((pos[b], Rs, com, Iw))
    return out


def _fk_at(m, q):
    o = Oracle(m, q.shape[0], precision=64)
    o.set("qpos", q)
    o.forward_kinematics()
    return _link_frames(m, o)


def _jacobians(m, q, h=1e-6):
    """Central-difference COM-velocity and angular-velocity Jacobians of every link, per env."""
    n, D = q.shape
    Jv = np.zeros((n, m.n_links, 3, D))
    Jw = np.zeros((n, m.n_links, 3, D))
    for i in range(D):
        dq = np.zeros_like(q)
        dq[:, i] = h
        fp, fm = _fk_at(m, q + dq), _fk_at(m, q - dq)
        for b in range(n):
            for l in range(m.n_links):
                Jv[b, l, :, i] = (fp[b][2][l] - fm[b][2][l]) / (2 * h)
                dR = (fp[b][1][l] - fm[b][1][l]) / (2 * h)
                W = dR @ fp[b][1][l].T * 0.5 + dR @ fm[b][1][l].T * 0.5  # skew(omega)
                Jw[b, l, :, i] = [W[2, 1], W[0, 2], W[1, 0]]
    return Jv, Jw


def test_mass_matrix_equals_kinetic_energy_hessian():
    """M(q) - diag(armature + damping dt) = sum_l m_l Jv^T Jv + Jw^T I_world Jw, with Jacobians by central
    differences of forward kinematics. A wrong inertia frame, COM offset or CRBA accumulation breaks it."""
    m, o, q, _ = _franka_state(3, 11)
    o.stage(STAGE_ID["crba"])
    D = m.n_dofs
    M = o.get("mass_mat").reshape(-1, D, D)
    A = m.arrays
    diag = A["dof_armature"] + A["dof_damping"] * m.options.dt
    Jv, Jw = _jacobians(m, q)
    frames = _fk_at(m, q)
    for b in range(q.shape[0]):
        Mfd = np.zeros((D, D))
        for l in range(m.n_links):
            Mfd += A["link_inertial_mass"][l] * Jv[b, l].T @ Jv[b, l] + Jw[b, l].T @ frames[b][3][l] @ Jw[b, l]
        err = np.abs(M[b] - np.diag(diag) - Mfd).max() / np.abs(Mfd).max()
        assert err < 1e-7, err


def test_ltdl_factor_reconstructs_mass_matrix_and_solves():
    """M = L^T diag(1/D_inv) L with unit lower-triangular L (reference test_mass_mat, test_rigid_physics.py:2315),
    and M acc_smooth = force."""
    m, o, _, _ = _franka_state(4, 3)
    for s in ("crba", "factor", "rne"):
        o.stage(STAGE_ID[s])
    D = m.n_dofs
    M = o.get("mass_mat").reshape(-1, D, D)
    L = o.get("mass_L").reshape(-1, D, D)
    Dinv = o.get("mass_D_inv")
    for b in range(M.shape[0]):
        Lb = np.tril(L[b])
        assert np.allclose(np.diag(Lb), 1.0)
        R = Lb.T @ np.diag(1.0 / Dinv[b]) @ Lb
        assert np.allclose(R, M[b], rtol=0, atol=1e-12 * np.abs(M[b]).max())
        assert np.allclose(M[b] @ o.get("acc_smooth")[b], o.get("force")[b], atol=1e-10 * np.abs(o.get("force")[b]).max())


def test_link_velocity_matches_finite_difference():
    """cd_vel is the velocity of the point at the subtree COM: origin velocity = cd_vel + cd_ang x (p - root_COM),
    which must equal d/dt of the FK origin along v; cd_ang equals the angular velocity of the link frame."""
    m, o, q, v = _franka_state(3, 5)
    h = 1e-6
    fp, fm = _fk_at(m, q + h * v), _fk_at(m, q - h * v)
    cdv = o.get("cd_vel").reshape(q.shape[0], -1, 3)
    cda = o.get("cd_ang").reshape(q.shape[0], -1, 3)
    com = o.get("root_COM").reshape(q.shape[0], -1, 3)
    pos = o.get("links_pos").reshape(q.shape[0], -1, 3)
    for b in range(q.shape[0]):
        for l in range(m.n_links):
            vfd = (fp[b][0][l] - fm[b][0][l]) / (2 * h)
            vp = cdv[b, l] + np.cross(cda[b, l], pos[b, l] - com[b, l])
            assert np.allclose(vp, vfd, atol=1e-6), (b, l)
            dR = (fp[b][1][l] - fm[b][1][l]) / (2 * h)
            W = dR @ (0.5 * (fp[b][1][l] + fm[b][1][l])).T
            assert np.allclose(cda[b, l], [W[2, 1], W[0, 2], W[1, 0]], atol=1e-6), (b, l)


def _potential(m, q):
    A = m.arrays
    frames = _fk_at(m, q)
    return np.array([sum(A["link_inertial_mass"][l] * G * fr[2][l][2] for l in range(m.n_links)) for fr in frames])


def test_gravity_bias_is_potential_gradient():
    """At v = 0 the bias force is dU/dq with U = sum m g z_com."""
    m, o, q, _ = _franka_state(3, 7)
    o.set("vel", np.zeros_like(q))
    o.forward_kinematics()
    for s in ("crba", "factor", "rne"):
        o.stage(STAGE_ID[s])
    bias = o.get("qf_bias")
    h = 1e-6
    for i in range(m.n_dofs):
        dq = np.zeros_like(q)
        dq[:, i] = h
        dU = (_potential(m, q + dq) - _potential(m, q - dq)) / (2 * h)
        assert np.allclose(bias[:, i], dU, atol=1e-6 * max(1.0, np.abs(dU).max())), i


def test_velocity_bias_is_lagrangian_coriolis_term():
    """With gravity off: bias = Mdot v - 1/2 d(v^T M v)/dq (Lagrange's equations), both by central differences
    of the oracle's own mass matrix along v and along each q_i."""
    m = franka_model(gravity=(0.0, 0.0, 0.0))
    n = 2
    rng = np.random.default_rng(9)
    q = franka_initial_qpos(m, n, 9)
    v = rng.normal(size=(n, m.n_dofs))

    def mass(qq):
        o = Oracle(m, n, precision=64)
        o.set("qpos", qq)
        o.set("ctrl_mode", np.full((n, m.n_dofs), 2.0))
        o.forward_kinematics()
        o.stage(STAGE_ID["crba"])
        return o.get("mass_mat").reshape(n, m.n_dofs, m.n_dofs)

    o = Oracle(m, n, precision=64)
    o.set("qpos", q)
    o.set("vel", v)
    o.set("ctrl_mode", np.full((n, m.n_dofs), 2.0))
    o.forward_kinematics()
    for s in ("crba", "factor", "rne"):
        o.stage(STAGE_ID[s])
    bias = o.get("qf_bias")
    h = 1e-5
    Mdot = (mass(q + h * v) - mass(q - h * v)) / (2 * h)
    want = np.einsum("bij,bj->bi", Mdot, v)
    for i in range(m.n_dofs):
        dq = np.zeros_like(q)
        dq[:, i] = h
        Mp, Mm = mass(q + dq), mass(q - dq)
        dE = (np.einsum("bi,bij,bj->b", v, Mp, v) - np.einsum("bi,bij,bj->b", v, Mm, v)) / (2 * h)
        want[:, i] -= 0.5 * dE
    assert np.allclose(bias, want, atol=1e-5 * max(1.0, np.abs(want).max()))


# ---------------------------------------------------------------------------------------------- collision
def _contacts(o, b=0):
    n = int(o.get("n_contacts")[b, 0])
    pos = o.get("contact_pos")[b, : 3 * n].reshape(n, 3)
    nrm = o.get("contact_normal")[b, : 3 * n].reshape(n, 3)
    pen = o.get("contact_penetration")[b, :n]
    return pos, nrm, pen


def _detect(m, o):
    o.forward_kinematics()
    for s in ("aabb", "broadphase", "narrowphase"):
        o.stage(STAGE_ID[s])


def test_sphere_on_plane_contact():
    r, z = 0.1, 0.08
    m = compile_model([plane_entity(), sphere_entity(r, pos=(0.3, -0.2, z))], SolverOptions())
    o = Oracle(m, 1, precision=64)
    _detect(m, o)
    pos, nrm, pen = _contacts(o)
    assert len(pen) == 1
    assert np.isclose(pen[0], r - z, atol=1e-9)
    assert np.allclose(np.abs(nrm[0]), [0, 0, 1], atol=1e-9)
    # the contact point is midway between the two surfaces
    assert np.allclose(pos[0], [0.3, -0.2, (z - r) / 2], atol=1e-9)


def test_resting_box_on_plane_gives_four_corner_contacts():
    s, d = np.array([0.2, 0.3, 0.1]), 1e-3
    m = compile_model([plane_entity(), box_entity(s, pos=(0, 0, s[2] / 2 - d))], SolverOptions())
    o = Oracle(m, 1, precision=64)
    _detect(m, o)
    pos, nrm, pen = _contacts(o)
    assert len(pen) == 4
    assert np.allclose(pen, d, atol=1e-9)
    corners = {(round(p[0], 6), round(p[1], 6)) for p in pos}
    assert corners == {(sx * s[0] / 2, sy * s[1] / 2) for sx in (-1, 1) for sy in (-1, 1)}


def test_crossed_horizontal_capsules_penetration():
    """Capsules along x and y, axes gap apart along z: depth r1 + r2 - gap along z."""
    r1, r2, gap = 0.05, 0.04, 0.07
    rot_x = np.array([np.cos(np.pi / 4), 0, np.sin(np.pi / 4), 0])   # axis z -> x
    rot_y = np.array([np.cos(np.pi / 4), -np.sin(np.pi / 4), 0, 0])  # axis z -> y
    m = compile_model([capsule_entity(r1, 0.4, pos=(0, 0, 1.0), quat=rot_x, name="a"),
                       capsule_entity(r2, 0.4, pos=(0, 0, 1.0 + gap), quat=rot_y, name="b")],
                      SolverOptions(gravity=(0.0, 0.0, 0.0)))
    o = Oracle(m, 1, precision=64)
    _detect(m, o)
    pos, nrm, pen = _contacts(o)
    assert len(pen) >= 1
    assert np.isclose(pen[0], r1 + r2 - gap, atol=1e-9)
    assert np.allclose(np.abs(nrm[0]), [0, 0, 1], atol=1e-9)


def test_box_box_mpr_penetration_depth():
    """Two axis-aligned boxes overlapping by delta along z: MPR finds depth delta along z (the shallowest axis)."""
    delta = 0.01
    m = compile_model([box_entity([0.4, 0.4, 0.2], pos=(0, 0, 1.0), name="a"),
                       box_entity([0.3, 0.3, 0.2], pos=(0.02, -0.01, 1.2 - delta), name="b")],
                      SolverOptions(gravity=(0.0, 0.0, 0.0)))
    o = Oracle(m, 1, precision=64)
    _detect(m, o)
    pos, nrm, pen = _contacts(o)
    assert len(pen) >= 1
    assert np.isclose(pen[0], delta, atol=1e-6)
    assert np.allclose(np.abs(nrm[0]), [0, 0, 1], atol=1e-6)


def test_sap_pairs_equal_brute_force_overlaps():
    """The broadphase pair set is exactly the set of valid pairs whose AABBs overlap (checked on random piles)."""
    m = box_pile_model()
    n = 6
    o = Oracle(m, n, precision=64)
    rng = np.random.default_rng(1)
    q = o.get("qpos")
    for b in range(n):  # squeeze the pile into a small region so that many AABBs overlap
        for k in range(20):
            q[b, 7 * k : 7 * k + 3] = [*rng.uniform(-0.12, 0.12, 2), rng.uniform(0.0, 0.4)]
    o.set("qpos", q)
    o.forward_kinematics()
    o.stage(STAGE_ID["aabb"])
    o.stage(STAGE_ID["broadphase"])
    lo = o.get("aabb_min").reshape(n, -1, 3)
    hi = o.get("aabb_max").reshape(n, -1, 3)
    pidx = m.arrays["collision_pair_idx"].reshape(m.n_geoms, m.n_geoms)
    for b in range(n):
        nb = int(o.get("n_broad")[b, 0])
        got = o.get("broad_pairs")[b, : 2 * nb].reshape(nb, 2).astype(int)
        assert len({tuple(p) for p in got}) == nb
        want = set()
        for i in range(m.n_geoms):
            for j in range(i + 1, m.n_geoms):
                if pidx[i, j] >= 0 and np.all(hi[b, i] > lo[b, j]) and np.all(lo[b, i] < hi[b, j]):
                    want.add((i, j))
        assert {tuple(p) for p in got} == want
        assert nb > 10


# ------------------------------------------------------------------------------------------------- solver
def test_newton_solution_satisfies_optimality_conditions():
    """At the solver's answer: M a - f_smooth - J^T lambda ~ 0, pyramid forces are >= 0 and vanish on inactive rows."""
    m = box_pile_model()
    n = 4
    o = Oracle(m, n, precision=64)
    o.step(80, 1)  # boxes land
    for s in ("crba", "factor", "rne", "aabb", "broadphase", "narrowphase", "constraints", "solve"):
        o.stage(STAGE_ID[s])
    D = m.n_dofs
    M = o.get("mass_mat").reshape(n, D, D)
    for b in range(n):
        nc = int(o.get("n_con")[b, 0])
        assert nc > 8
        J = o.get("jac")[b, : nc * D].reshape(nc, D)
        lam = o.get("efc_force")[b, :nc]
        r = M[b] @ o.get("qacc")[b] - o.get("qf_smooth")[b] - J.T @ lam
        scale = np.abs(o.get("qf_smooth")[b]).max() + np.abs(J.T @ lam).max()
        assert np.abs(r).max() < 1e-4 * scale
        ne, nf = int(o.get("n_con_eq")[b, 0]), int(o.get("n_con_fl")[b, 0])
        contact = lam[ne + nf :]
        assert contact.min() > -1e-12
        active = o.get("active")[b, :nc].astype(bool)[ne + nf :]
        assert np.all(contact[~active] == 0)


# ------------------------------------------------------------------------------- reference known answers
BOX_BOX = """<mujoco model="two_boxes"><option timestep="0.01"/>
<default><geom contype="1" conaffinity="1" condim="3" friction="1. 0.5 0.5"/></default>
<worldbody>
  <body name="box1" pos="0 0 0.2"><geom type="box" size="0.2 0.2 0.2"/><joint name="root1" type="free"/></body>
  <body name="box2" pos="0 0 0.8"><geom type="box" size="0.2 0.2 0.2"/><joint name="root2" type="free"/></body>
</worldbody></mujoco>"""


@pytest.mark.parametrize("precision", [64, 32])
def test_reference_box_box_stack_settles(tmp_path, precision):
    """test_box_box_dynamics (tests/test_rigid_physics.py:1032-1053): a 0.4 m cube dropped with a random tilt onto
    another settles with its centre at z = 0.6 +- 2e-3 and |v| < 1e-2 after 100 steps."""
    ent = load_mjcf(_write(tmp_path, "bb.xml", BOX_BOX), "boxes")
    m = compile_model([plane_entity(friction=1.0), ent], SolverOptions(dt=0.01))  # the MJCF floor plane
    n = 6
    o = Oracle(m, n, precision=precision)
    rng = np.random.default_rng(4)
    q = np.zeros((n, 14))
    for b in range(n):
        e = np.array([*(0.15 * rng.random(2)), np.pi * rng.random()])
        cz, sz = np.cos(e / 2), np.sin(e / 2)
        qx = np.array([cz[0], sz[0], 0, 0])
        qy = np.array([cz[1], 0, sz[1], 0])
        qz = np.array([cz[2], 0, 0, sz[2]])
        q[b] = [0, 0, 0.2, 1, 0, 0, 0, 0, 0, 0.65 + 0.1 * rng.random(), *_quat_mul(_quat_mul(qz, qy), qx)]
    o.set("qpos", q)
    o.forward_kinematics()
    o.step(101, 1)
    for _ in range(9):
        o.step(1, 1)
        assert np.abs(o.get("vel")).max() < 1e-2
    assert np.allclose(o.get("qpos")[:, 9], 0.6, atol=2e-3)
