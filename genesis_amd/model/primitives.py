"""Entity specs for primitive morphs (Plane, Box, Sphere, Capsule).

Follows ``rigid_entity.py:282-378`` ``_load_primitive``: one link, a free joint (or none when fixed), one collision
geom whose ``data`` holds the primitive parameters; inertia from the shape and the material density, with the link
frame at the centre of mass (``rigid_entity.py:958-1047`` ``_align_link``; a centred primitive needs no rotation).
The plane's collision mesh is a 1 km x 1 km x 1 cm slab below the surface (``genesis/utils/mesh.py:946-983``).
"""

from __future__ import annotations

import numpy as np

from . import meshes as mu
from . import transforms as tf
from .specs import (
    GEOM_BOX,
    GEOM_CAPSULE,
    GEOM_PLANE,
    GEOM_SPHERE,
    JOINT_FREE,
    EntitySpec,
    GeomSpec,
    JointSpec,
    LinkSpec,
    default_sol_params,
)


def plane_entity(normal=(0.0, 0.0, 1.0), plane_size=(1e3, 1e3), friction=1.0, name="plane") -> EntitySpec:
    normal = np.asarray(normal, dtype=np.float64)
    normal = normal / np.linalg.norm(normal)
    thickness = 1e-2
    verts, faces = mu.box_mesh([plane_size[0], plane_size[1], thickness])
    verts = verts.copy()
    verts[:, 2] -= thickness / 2
    z = np.array([0.0, 0.0, 1.0])
    if not np.allclose(normal, z):
        c = np.cross(z, normal)
        if np.linalg.norm(c) < 1e-12:
            R = np.diag([1.0, -1.0, -1.0])
        else:
            R = tf.quat_to_R(tf.axis_angle_quat(c, np.arctan2(np.linalg.norm(c), normal @ z)))
        verts = verts @ R.T
    _, com, _ = mu.mass_properties(verts, faces)
    geom = GeomSpec(
        type=GEOM_PLANE, pos=np.zeros(3), quat=np.array([1.0, 0, 0, 0]), data=normal.copy(), verts=verts,
        faces=faces, friction=friction, sol_params=default_sol_params(), center=com, name=name,
    )
    link = LinkSpec(
        name=f"{name}_baselink", parent=-1, pos=np.zeros(3), quat=np.array([1.0, 0, 0, 0]),
        inertial_pos=np.zeros(3), inertial_quat=np.array([1.0, 0, 0, 0]), inertial_i=np.zeros((3, 3)),
        inertial_mass=0.0, joints=[], geoms=[geom],
    )
    return EntitySpec(name=name, links=[link])


def _free_body(name, pos, quat, mass, inertia, geom, fixed=False) -> EntitySpec:
    pos = np.asarray(pos, dtype=np.float64)
    quat = tf.quat_normalize(quat)
    joints = []
    if not fixed:
        joints.append(JointSpec(name=f"{name}_baselink_joint", type=JOINT_FREE, init_qpos=np.concatenate([pos, quat])))
    link = LinkSpec(
        name=f"{name}_baselink", parent=-1, pos=pos, quat=quat, inertial_pos=np.zeros(3),
        inertial_quat=np.array([1.0, 0, 0, 0]), inertial_i=inertia, inertial_mass=mass, joints=joints, geoms=[geom],
    )
    return EntitySpec(name=name, links=[link])


def box_entity(size, pos=(0, 0, 0), quat=(1, 0, 0, 0), rho=1000.0, friction=1.0, fixed=False, name="box"):
    size = np.asarray(size, dtype=np.float64)
    verts, faces = mu.box_mesh(size)
    geom = GeomSpec(
        type=GEOM_BOX, pos=np.zeros(3), quat=np.array([1.0, 0, 0, 0]), data=size.copy(), verts=verts, faces=faces,
        friction=friction, sol_params=default_sol_params(), center=np.zeros(3), name=name,
    )
    m = rho * float(np.prod(size))
    x, y, z = size
    inertia = np.diag([m / 12 * (y * y + z * z), m / 12 * (x * x + z * z), m / 12 * (x * x + y * y)])
    return _free_body(name, pos, quat, m, inertia, geom, fixed)


def sphere_entity(radius, pos=(0, 0, 0), quat=(1, 0, 0, 0), rho=1000.0, friction=1.0, fixed=False, name="sphere"):
    geom = GeomSpec(
        type=GEOM_SPHERE, pos=np.zeros(3), quat=np.array([1.0, 0, 0, 0]), data=np.array([radius]),
        verts=mu.sphere_points(radius), friction=friction, sol_params=default_sol_params(), center=np.zeros(3),
        name=name,
    )
    m = rho * 4.0 / 3.0 * np.pi * radius**3
    return _free_body(name, pos, quat, m, np.eye(3) * 0.4 * m * radius**2, geom, fixed)


def capsule_entity(radius, length, pos=(0, 0, 0), quat=(1, 0, 0, 0), rho=1000.0, friction=1.0, fixed=False,
                   name="capsule"):
    geom = GeomSpec(
        type=GEOM_CAPSULE, pos=np.zeros(3), quat=np.array([1.0, 0, 0, 0]), data=np.array([radius, length]),
        verts=mu.capsule_points(radius, length), friction=friction, sol_params=default_sol_params(),
        center=np.zeros(3), name=name,
    )
    m_cyl = rho * np.pi * radius**2 * length
    m_sph = rho * 4.0 / 3.0 * np.pi * radius**3
    m = m_cyl + m_sph
    izz = 0.5 * m_cyl * radius**2 + 0.4 * m_sph * radius**2
    ixx = m_cyl * (3 * radius**2 + length**2) / 12 + m_sph * (0.4 * radius**2 + 0.25 * length**2 + 3 / 8 * radius * length)
    return _free_body(name, pos, quat, m, np.diag([ixx, ixx, izz]), geom, fixed)


def hull_entity(points, pos=(0, 0, 0), quat=(1, 0, 0, 0), rho=1000.0, friction=1.0, fixed=False, name="hull"):
    """A rigid body whose collision geom is the convex hull of ``points`` (given about the body origin), with the
    link frame at the hull's centre of mass (``rigid_entity.py:925-1007`` convexify, ``:958-1047`` ``_align_link``)."""
    from .specs import GEOM_MESH

    verts, faces = mu.convex_hull(np.asarray(points, dtype=np.float64))
    vol, com, inertia = mu.mass_properties(verts, faces)
    verts = verts - com
    pos = np.asarray(pos, dtype=np.float64) + tf.quat_to_R(tf.quat_normalize(quat)) @ com
    geom = GeomSpec(
        type=GEOM_MESH, pos=np.zeros(3), quat=np.array([1.0, 0, 0, 0]), data=np.zeros(7), verts=verts, faces=faces,
        friction=friction, sol_params=default_sol_params(), center=np.zeros(3), name=name,
    )
    return _free_body(name, pos, quat, rho * vol, rho * inertia, geom, fixed)
