"""Enumerations stored in solver arrays; values equal the reference's (genesis/constants.py:17-129)."""

import enum


class IntEnum(enum.IntEnum):
    def __repr__(self):
        return f"<gs.{self.__class__.__name__}.{self.name}: {self.value}>"


class GEOM_TYPE(IntEnum):
    PLANE = 0
    SPHERE = 1
    ELLIPSOID = 2
    CYLINDER = 3
    CAPSULE = 4
    BOX = 5
    MESH = 6
    TERRAIN = 7


class JOINT_TYPE(IntEnum):
    FIXED = 0
    REVOLUTE = 1
    PRISMATIC = 2
    SPHERICAL = 3
    FREE = 4


class EQUALITY_TYPE(IntEnum):
    CONNECT = 0
    WELD = 1
    JOINT = 2


class CTRL_MODE(IntEnum):
    POSITION = 0
    VELOCITY = 1
    FORCE = 2


class integrator(IntEnum):
    Euler = 0
    implicitfast = 1
    approximate_implicitfast = 2


class constraint_solver(IntEnum):
    CG = 0
    Newton = 1


class broadphase_traversal(IntEnum):
    SAP = 0
    ALL_VS_ALL = 1


class backend(IntEnum):
    cpu = 0
    gpu = 1
    cuda = 2
    amdgpu = 3
    metal = 4
