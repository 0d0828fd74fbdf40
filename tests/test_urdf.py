"""URDF subset loader (genesis_amd/model/urdf.py), after the reference's URDF path (utils/mjcf.py:62-198 build_model
under MuJoCo's URDF compile, utils/urdf.py:131-400 parse_urdf for force ranges, armature and gains, urdf.py:436-660
merge_fixed_links, rigid_entity.py:653-690 the floating base). MuJoCo is absent, so the loader's results are pinned by
what the mathematics fixes:

- URDF's rpy convention against scipy's extrinsic x-y-z rotation;
- forward kinematics of the compiled model (FP64 restatement) at random joint positions against the frames chained
  by hand from the URDF with scipy rotations (joint origins, axes, hinge and slide motion);
- each link's mass and world-frame inertia about its centre of mass against the URDF's, rotated by scipy;
- merging fixed links changes nothing physical: the same joint-space mass matrix and the same trajectory as the
  unmerged model;
- the reference's joint defaults (armature 0.1, kp 100 / kv 10 or the safety controller's gains, effort as force
  range, dynamics as damping / friction loss, continuous joints unlimited), the floating base, mesh scale.
Parity with MuJoCo's own URDF compile itself is unpinned (no MuJoCo here).
"""

from __future__ import annotations

import os

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from genesis_amd.model.compiler import SolverOptions, compile_model
from genesis_amd.model.urdf import load_urdf, rpy_to_R
from genesis_amd.runtime import STAGE_ID
from tests.oracle_lib import Oracle

ARM = """<robot name="arm">
  <link name="base">
    <inertial><origin xyz="0 0 0.05" rpy="0 0 0"/><mass value="2.0"/>
      <inertia ixx="0.02" ixy="0" ixz="0" iyy="0.03" iyz="0" izz="0.04"/></inertial>
    <collision><origin xyz="0 0 0.05"/><geometry><box size="0.3 0.2 0.1"/></geometry></collision>
  </link>
  <link name="rail">
    <inertial><origin xyz="0.02 0 0" rpy="0.1 -0.2 0.3"/><mass value="0.5"/>
      <inertia ixx="0.004" ixy="0.0003" ixz="-0.0002" iyy="0.005" iyz="0.0001" izz="0.003"/></inertial>
    <collision><origin xyz="0 0 0"/><geometry><sphere radius="0.04"/></geometry></collision>
  </link>
  <link name="upper">
    <inertial><origin xyz="0 0 0.2" rpy="0.3 0.2 0.1"/><mass value="0.8"/>
      <inertia ixx="0.011" ixy="0.001" ixz="0.0005" iyy="0.012" iyz="0.0002" izz="0.002"/></inertial>
    <collision><origin xyz="0 0 0.2" rpy="0 0.1 0"/><geometry><box size="0.05 0.05 0.4"/></geometry></collision>
  </link>
  <link name="lower">
    <inertial><origin xyz="0 0 0.15" rpy="0 0 0"/><mass value="0.6"/>
      <inertia ixx="0.005" ixy="0" ixz="0" iyy="0.005" iyz="0" izz="0.001"/></inertial>
    <collision><origin xyz="0 0 0.3"/><geometry><sphere radius="0.03"/></geometry></collision>
  </link>
  <link name="tool">
    <inertial><origin xyz="0.01 0 0.02" rpy="0 0.4 0"/><mass value="0.3"/>
      <inertia ixx="0.0006" ixy="0.00005" ixz="0" iyy="0.0008" iyz="0" izz="0.0004"/></inertial>
    <collision><origin xyz="0 0 0.02" rpy="0.2 0 0"/><geometry><box size="0.06 0.04 0.04"/></geometry></collision>
  </link>
  <link name="marker"/>
  <joint name="slide" type="prismatic"><parent link="base"/><child link="rail"/>
    <origin xyz="0.1 0 0.12" rpy="0 0 0.4"/><axis xyz="1 0 0"/>
    <limit lower="-0.2" upper="0.2" effort="80" velocity="1"/></joint>
  <joint name="shoulder" type="revolute"><parent link="rail"/><child link="upper"/>
    <origin xyz="0 0 0.05" rpy="0 0 1.5707963"/><axis xyz="0 1 0"/>
    <limit lower="-1.5" upper="1.5" effort="30" velocity="2"/><dynamics damping="0.5" friction="0.1"/></joint>
  <joint name="elbow" type="continuous"><parent link="upper"/><child link="lower"/>
    <origin xyz="0 0 0.4" rpy="0.2 0 0"/><axis xyz="1 0 1"/>
    <safety_controller k_position="50" k_velocity="5"/></joint>
  <joint name="wrist_fixed" type="fixed"><parent link="lower"/><child link="tool"/>
    <origin xyz="0 0.05 0.3" rpy="0 0.5 -0.3"/></joint>
  <joint name="marker_fixed" type="fixed"><parent link="tool"/><child link="marker"/>
    <origin xyz="0 0 0.1"/></joint>
</robot>
"""


@pytest.fixture()
def arm_path(tmp_path):
    p = os.path.join(str(tmp_path), "arm.urdf")
    with open(p, "w") as f:
        f.write(ARM)
    return p


def T(xyz=(0, 0, 0), rpy=(0, 0, 0)):
    out = np.eye(4)
    out[:3, :3] = Rotation.from_euler("xyz", rpy).as_matrix()  # scipy lower case: extrinsic axes
    out[:3, 3] = xyz
    return out


def _hinge(axis, q):
    out = np.eye(4)
    out[:3, :3] = Rotation.from_rotvec(np.asarray(axis) / np.linalg.norm(axis) * q).as_matrix()
    return out


def _slide(axis, q):
    out = np.eye(4)
    out[:3, 3] = np.asarray(axis) / np.linalg.norm(axis) * q
    return out


def _expected_frames(q, base=np.eye(4)):
    """World frames of base, rail, upper, lower, tool (merged into lower) chained from the URDF text above."""
    base_T = base
    rail_T = base_T @ T((0.1, 0, 0.12), (0, 0, 0.4)) @ _slide((1, 0, 0), q[0])
    upper_T = rail_T @ T((0, 0, 0.05), (0, 0, 1.5707963)) @ _hinge((0, 1, 0), q[1])
    lower_T = upper_T @ T((0, 0, 0.4), (0.2, 0, 0)) @ _hinge((1, 0, 1), q[2])
    tool_T = lower_T @ T((0, 0.05, 0.3), (0, 0.5, -0.3))
    return {"base": base_T, "rail": rail_T, "upper": upper_T, "lower": lower_T, "tool": tool_T}


def test_rpy_is_extrinsic_xyz():
    rng = np.random.default_rng(0)
    for _ in range(20):
        rpy = rng.uniform(-np.pi, np.pi, 3)
        assert np.allclose(rpy_to_R(rpy), Rotation.from_euler("xyz", rpy).as_matrix(), atol=1e-12)
    # yaw of 90 degrees takes x to y
    assert np.allclose(rpy_to_R((0, 0, np.pi / 2)) @ [1, 0, 0], [0, 1, 0], atol=1e-12)


def _model(path, **kw):
    ent = load_urdf(path, **kw)
    return ent, compile_model([ent], SolverOptions(gravity=(0.0, 0.0, -9.81)))


def test_structure_and_joint_defaults(arm_path):
    ent, m = _model(arm_path, fixed=True)
    names = [l.name for l in ent.links]
    assert names == ["base", "rail", "upper", "lower"]  # tool and marker merged into lower
    joints = {j.name: j for l in ent.links for j in l.joints}
    assert set(joints) == {"slide", "shoulder", "elbow"}
    assert joints["slide"].limit == (-0.2, 0.2) and joints["slide"].force_range == (-80.0, 80.0)
    assert joints["shoulder"].limit == (-1.5, 1.5) and joints["shoulder"].damping == 0.5
    assert joints["shoulder"].frictionloss == 0.1
    assert joints["elbow"].limit is None and joints["elbow"].force_range == (-np.inf, np.inf)
    assert all(j.armature == 0.1 for j in joints.values())
    assert joints["slide"].act_gain == 100.0 and joints["slide"].act_bias == (0.0, -100.0, -10.0)
    assert joints["elbow"].act_gain == 50.0 and joints["elbow"].act_bias == (0.0, -50.0, -5.0)
    assert m.n_dofs == 3
    # floating base: a free root joint in front of the three
    ent_f, m_f = _model(arm_path, pos=(0.3, -0.2, 1.0))
    assert m_f.n_dofs == 9 and ent_f.links[0].joints[0].type == 4  # JOINT_FREE
    assert ent_f.links[0].joints[0].armature == 0.0
    assert np.allclose(m_f.arrays["qpos0"][:3], [0.3, -0.2, 1.0])
    # no fixed-link merging: all six links stay, and a link without <inertial> has no mass
    ent_u, _ = _model(arm_path, fixed=True, merge_fixed=False)
    assert [l.name for l in ent_u.links] == ["base", "rail", "upper", "lower", "tool", "marker"]
    assert ent_u.links[-1].inertial_mass == 0.0
    # links_to_keep holds a fixed link back
    ent_k, _ = _model(arm_path, fixed=True, links_to_keep=("tool",))
    assert [l.name for l in ent_k.links] == ["base", "rail", "upper", "lower", "tool"]


def _fk(m, q):
    o = Oracle(m, 1, precision=64)
    o.set("qpos", np.asarray(q, dtype=np.float64)[None, :])
    o.forward_kinematics()
    return o.get("links_pos")[0].reshape(-1, 3), o.get("links_quat")[0].reshape(-1, 4)


def _R_of(quat_wxyz):
    w, x, y, z = quat_wxyz
    return Rotation.from_quat([x, y, z, w]).as_matrix()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_forward_kinematics_matches_the_urdf_frames(arm_path, seed):
    ent, m = _model(arm_path, fixed=True)
    q = np.random.default_rng(seed).uniform(-1.0, 1.0, 3) * [0.2, 1.4, 3.0]
    pos, quat = _fk(m, q)
    want = _expected_frames(q)
    for i, l in enumerate(ent.links):
        assert np.allclose(pos[i], want[l.name][:3, 3], atol=1e-9), l.name
        assert np.allclose(_R_of(quat[i]), want[l.name][:3, :3], atol=1e-9), l.name


def _urdf_inertial(xyz, rpy, mass, I6):
    ixx, ixy, ixz, iyy, iyz, izz = I6
    I = np.array([[ixx, ixy, ixz], [ixy, iyy, iyz], [ixz, iyz, izz]])
    Tin = T(xyz, rpy)
    return mass, Tin, Tin[:3, :3] @ I @ Tin[:3, :3].T


def test_link_mass_and_inertia_match_the_urdf(arm_path):
    """Mass, world centre of mass and world inertia about it, at a random configuration: the lower link carries the
    merged tool (parallel-axis sum written out here in the world frame)."""
    ent, m = _model(arm_path, fixed=True)
    q = np.array([0.05, 0.7, -1.1])
    frames = _expected_frames(q)
    spec = {"base": ((0, 0, 0.05), (0, 0, 0), 2.0, (0.02, 0, 0, 0.03, 0, 0.04)),
            "rail": ((0.02, 0, 0), (0.1, -0.2, 0.3), 0.5, (0.004, 0.0003, -0.0002, 0.005, 0.0001, 0.003)),
            "upper": ((0, 0, 0.2), (0.3, 0.2, 0.1), 0.8, (0.011, 0.001, 0.0005, 0.012, 0.0002, 0.002)),
            "lower": ((0, 0, 0.15), (0, 0, 0), 0.6, (0.005, 0, 0, 0.005, 0, 0.001)),
            "tool": ((0.01, 0, 0.02), (0, 0.4, 0), 0.3, (0.0006, 0.00005, 0, 0.0008, 0, 0.0004))}

    def world(name):
        mass, Tin, I = _urdf_inertial(*spec[name][:2], spec[name][2], spec[name][3])
        Fw = frames[name]
        com = (Fw @ Tin)[:3, 3]
        Rw = Fw[:3, :3]
        return mass, com, Rw @ I @ Rw.T

    def combine(parts):
        M = sum(p[0] for p in parts)
        c = sum(p[0] * p[1] for p in parts) / M
        I = sum(p[2] + p[0] * (np.dot(c - p[1], c - p[1]) * np.eye(3) - np.outer(c - p[1], c - p[1])) for p in parts)
        return M, c, I

    want = {n: world(n) for n in ("base", "rail", "upper")}
    want["lower"] = combine([world("lower"), world("tool")])
    pos, quat = _fk(m, q)
    A = m.arrays
    for i, l in enumerate(ent.links):
        Rl = _R_of(quat[i])
        com = pos[i] + Rl @ A["link_inertial_pos"][i]
        Ri = Rl @ _R_of(A["link_inertial_quat"][i])
        I = Ri @ np.asarray(A["link_inertial_i"][i]).reshape(3, 3) @ Ri.T
        M, c, Iw = want[l.name]
        assert abs(A["link_inertial_mass"][i] - M) < 1e-12, l.name
        assert np.allclose(com, c, atol=1e-9), l.name
        assert np.allclose(I, Iw, atol=1e-9), l.name


def test_merging_fixed_links_keeps_the_dynamics(arm_path):
    """Merged and unmerged models: the same joint-space mass matrix at random configurations, and the same
    200-substep trajectory of the floating arm under gravity and PD targets with collisions off (merging moves the
    tool's box onto the lower link, which changes which geom pairs are filtered as neighbours)."""
    opt = SolverOptions(enable_collision=False)
    mm = compile_model([load_urdf(arm_path)], opt)
    mu_ = compile_model([load_urdf(arm_path, merge_fixed=False)], opt)
    assert mm.n_dofs == mu_.n_dofs == 9 and mm.n_links < mu_.n_links
    rng = np.random.default_rng(4)
    B = 4
    q = np.tile(mm.arrays["qpos0"], (B, 1))
    q[:, 3:7] = Rotation.from_rotvec(rng.normal(0, 0.5, (B, 3))).as_quat()[:, [3, 0, 1, 2]]
    q[:, 7:] = rng.uniform(-0.5, 0.5, (B, 3)) * [0.3, 1.0, 1.0]
    outs = []
    for m in (mm, mu_):
        o = Oracle(m, B, precision=64)
        o.set("qpos", q)
        o.forward_kinematics()
        o.stage(STAGE_ID["crba"])
        outs.append(o)
    M0 = outs[0].get("mass_mat")
    assert np.abs(M0).max() > 0.1
    assert np.allclose(M0, outs[1].get("mass_mat"), rtol=1e-12, atol=1e-12)
    for o in outs:
        o.set("ctrl_mode", np.zeros((B, 9)))
        o.set("ctrl_pos", np.tile([0, 0, 0, 0, 0, 0, 0.1, 0.5, -0.8], (B, 1)))
        o.step(200)
    d = np.abs(outs[0].get("qpos") - outs[1].get("qpos")).max()
    assert d < 1e-9, d


def test_mesh_collision_with_scale(tmp_path):
    obj = os.path.join(str(tmp_path), "cube.obj")
    v = [(x, y, z) for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.5, 0.5)]
    faces = [(1, 3, 4, 2), (5, 6, 8, 7), (1, 2, 6, 5), (3, 7, 8, 4), (1, 5, 7, 3), (2, 4, 8, 6)]
    with open(obj, "w") as f:
        f.writelines(f"v {a} {b} {c}\n" for a, b, c in v)
        f.writelines("f " + " ".join(map(str, fc)) + "\n" for fc in faces)
    urdf = os.path.join(str(tmp_path), "m.urdf")
    with open(urdf, "w") as f:
        f.write('<robot name="m"><link name="a"><inertial><mass value="1"/><inertia ixx="0.1" iyy="0.1" izz="0.1"/>'
                '</inertial><collision><geometry><mesh filename="cube.obj" scale="0.2 0.4 0.6"/></geometry></collision>'
                '</link></robot>')
    ent = load_urdf(urdf)
    g = ent.links[0].geoms[0]
    ext = np.sort(g.verts.max(axis=0) - g.verts.min(axis=0))
    assert np.allclose(ext, [0.2, 0.4, 0.6], atol=1e-9)


@pytest.mark.parametrize("bad,exc", [('<geometry><cylinder radius="0.1" length="0.2"/></geometry>', NotImplementedError)])
def test_unsupported_inputs_raise(tmp_path, bad, exc):
    urdf = os.path.join(str(tmp_path), "bad.urdf")
    with open(urdf, "w") as f:
        f.write(f'<robot name="b"><link name="a"><collision>{bad}</collision></link></robot>')
    with pytest.raises(exc):
        load_urdf(urdf)
    with open(urdf, "w") as f:
        f.write('<robot name="b"><link name="a"/><link name="c"/><joint name="j" type="planar"><parent link="a"/>'
                '<child link="c"/></joint></robot>')
    with pytest.raises(NotImplementedError):
        load_urdf(urdf)
