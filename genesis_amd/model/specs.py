"""Plain-data description of a rigid entity before it is compiled into solver arrays.

The fields mirror the ``l_info`` / ``j_info`` / ``g_info`` / ``eq_info`` dictionaries the reference builds while
loading a morph (``genesis/utils/mjcf.py:229-404`` ``parse_link``, ``:441-615`` ``parse_geom``;
``genesis/engine/entities/rigid_entity/rigid_entity.py:2205-2299`` ``_add_by_info``).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# Geometry / joint / equality / control codes. Values are stored in solver arrays, so they equal the
# reference enums bit for bit (genesis/constants.py:17-49).
GEOM_PLANE, GEOM_SPHERE, GEOM_ELLIPSOID, GEOM_CYLINDER, GEOM_CAPSULE, GEOM_BOX, GEOM_MESH, GEOM_TERRAIN = range(8)
JOINT_FIXED, JOINT_REVOLUTE, JOINT_PRISMATIC, JOINT_SPHERICAL, JOINT_FREE = range(5)
EQ_CONNECT, EQ_WELD, EQ_JOINT = range(3)
CTRL_POSITION, CTRL_VELOCITY, CTRL_FORCE = range(3)

# MuJoCo default impedance parameters (solimp) and damping ratio, appended to a time constant that the solver
# replaces by its default when it is below machine epsilon (genesis/utils/mjcf.py:118-131).
DEFAULT_SOLIMP = (0.9, 0.95, 0.001, 0.5, 2.0)
MIN_TIMECONST = float(np.finfo(np.float64).eps)


def default_sol_params():
    return np.array([MIN_TIMECONST, 1.0, *DEFAULT_SOLIMP], dtype=np.float64)


@dataclass
class JointSpec:
    name: str
    type: int
    pos: np.ndarray = field(default_factory=lambda: np.zeros(3))
    axis: np.ndarray = field(default_factory=lambda: np.array([0.0, 0.0, 1.0]))
    limit: tuple | None = None  # (lo, hi) or None for unlimited
    armature: float = 0.0
    damping: float = 0.0
    stiffness: float = 0.0
    frictionloss: float = 0.0
    sol_params: np.ndarray = field(default_factory=default_sol_params)
    act_gain: float = 0.0
    act_bias: tuple = (0.0, 0.0, 0.0)
    force_range: tuple = (-np.inf, np.inf)
    init_qpos: np.ndarray | None = None

    @property
    def n_qs(self):
        return {JOINT_FIXED: 0, JOINT_REVOLUTE: 1, JOINT_PRISMATIC: 1, JOINT_SPHERICAL: 4, JOINT_FREE: 7}[self.type]

    @property
    def n_dofs(self):
        return {JOINT_FIXED: 0, JOINT_REVOLUTE: 1, JOINT_PRISMATIC: 1, JOINT_SPHERICAL: 3, JOINT_FREE: 6}[self.type]


@dataclass
class GeomSpec:
    type: int
    pos: np.ndarray
    quat: np.ndarray
    data: np.ndarray  # type-specific parameters (box: full extents, sphere: radius, capsule: radius, length, plane: normal)
    verts: np.ndarray  # collision vertices in the geom frame (convex hull for meshes)
    faces: np.ndarray | None = None
    friction: float = 1.0
    sol_params: np.ndarray = field(default_factory=default_sol_params)
    contype: int = 1
    conaffinity: int = 1
    center: np.ndarray | None = None  # centre of mass of the collision shape in the geom frame
    name: str = ""


@dataclass
class LinkSpec:
    name: str
    parent: int  # index within the entity, -1 for the world
    pos: np.ndarray
    quat: np.ndarray
    inertial_pos: np.ndarray
    inertial_quat: np.ndarray
    inertial_i: np.ndarray  # 3x3
    inertial_mass: float
    joints: list = field(default_factory=list)
    geoms: list = field(default_factory=list)
    invweight: np.ndarray | None = None


@dataclass
class EqualitySpec:
    """An equality constraint (solver.py:689-938). JOINT: obj1/obj2 are joint names and data[0:5] the polynomial
    coefficients. CONNECT: obj1/obj2 are link names, data[0:3] the anchor in obj1's frame and data[3:6] the same point
    in obj2's frame. WELD: data[0:3] the anchor in obj2's frame, data[3:6] in obj1's frame, data[6:10] the pose of
    obj2 relative to obj1 (quaternion), data[10] the torque scale. at_qpos0: fill data[3:6] (and, for a weld whose
    data[6:10] is zero, the relative pose) from the configuration at qpos0, so that the constraint holds there (what
    MuJoCo's model compiler does for MJCF equalities)."""

    type: int
    obj1: str
    obj2: str
    data: np.ndarray  # 11 floats
    sol_params: np.ndarray
    at_qpos0: bool = False


@dataclass
class EntitySpec:
    name: str
    links: list
    equalities: list = field(default_factory=list)
    is_local_collision_mask: bool = False
    gravity_compensation: float = 0.0
