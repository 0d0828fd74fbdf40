"""URDF subset loader.

The reference compiles a URDF file with MuJoCo (``genesis/utils/mjcf.py:62-198`` ``build_model``: the URDF tree with a
``<mujoco><compiler fusestatic="false" inertiafromgeom="false" autolimits="true" .../>`` extension, solver time
constants forced to their near-zero "use the default" value, links without ``<inertial>`` given zero mass and inertia)
and reads it like an MJCF model; its own URDF parser (``genesis/utils/urdf.py:131-400`` ``parse_urdf``) then supplies
what MuJoCo does not parse: force ranges from ``<limit effort>``, the default rotor armature, and the PD gains
(``rigid_entity.py:508-560``). Before that, links hanging on fixed joints are merged into their parents unless listed
in ``links_to_keep`` (``urdf.py:436-472`` ``merge_fixed_links``, ``:572-660``), and a root link attached to nothing
gets a free joint unless the morph is ``fixed`` (``rigid_entity.py:653-690``).

MuJoCo is not part of this build: this module parses the URDF subset the rigid step needs (links with inertials and
collision geometry ``box`` / ``sphere`` / ``mesh``, joints ``revolute`` / ``continuous`` / ``prismatic`` / ``fixed`` /
``floating`` with limits, dynamics and safety controllers), merges fixed links as the reference does, and writes the
result as the ``<mujoco>`` tree MuJoCo would build from it, which ``mjcf.MJCFLoader`` compiles (inertial principal
axes, mesh frames, breadth-first link order). Visual geometry is not loaded (it does not enter the step).
"""

from __future__ import annotations

import copy
import os
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field

import numpy as np

from . import transforms as tf
from .mjcf import MJCFLoader
from .specs import JOINT_FREE, JOINT_PRISMATIC, JOINT_REVOLUTE, EntitySpec

DEFAULT_ARMATURE = 0.1  # gs.morphs.URDF default_armature (options/morphs.py:1026)
DEFAULT_KP, DEFAULT_KV = 100.0, 10.0  # utils/geom.py:2116-2121 default_dofs_kp / default_dofs_kv


def _vec(s, n=3, default=0.0):
    if s is None:
        return np.full(n, default, dtype=np.float64)
    v = np.array([float(x) for x in s.split()], dtype=np.float64)
    if v.size != n:
        raise ValueError(f"URDF: expected {n} numbers, got {s!r}")
    return v


def rpy_to_R(rpy) -> np.ndarray:
    """URDF ``rpy``: fixed-axis roll about x, then pitch about y, then yaw about z (R = Rz Ry Rx)."""
    r, p, y = rpy
    cr, sr, cp, sp, cy, sy = np.cos(r), np.sin(r), np.cos(p), np.sin(p), np.cos(y), np.sin(y)
    Rx = np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])
    Ry = np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]])
    Rz = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def _origin(elem) -> np.ndarray:
    T = np.eye(4)
    o = elem.find("origin") if elem is not None else None
    if o is not None:
        T[:3, :3] = rpy_to_R(_vec(o.attrib.get("rpy")))
        T[:3, 3] = _vec(o.attrib.get("xyz"))
    return T


@dataclass
class _Inertial:
    mass: float
    origin: np.ndarray  # 4x4, inertial frame in the link frame
    inertia: np.ndarray  # 3x3 in the inertial frame


@dataclass
class _Collision:
    origin: np.ndarray
    kind: str  # box | sphere | mesh
    size: np.ndarray  # box: full extents; sphere: radius; mesh: scale
    filename: str = ""


@dataclass
class _Link:
    name: str
    inertial: _Inertial | None
    collisions: list = field(default_factory=list)


@dataclass
class _Joint:
    name: str
    type: str
    parent: str
    child: str
    origin: np.ndarray
    axis: np.ndarray
    lower: float | None = None
    upper: float | None = None
    effort: float | None = None
    damping: float = 0.0
    friction: float = 0.0
    k_position: float | None = None
    k_velocity: float | None = None


def _parse(path: str):
    root = ET.parse(path).getroot()
    if root.tag != "robot":
        raise ValueError(f"{path}: not a URDF file (root element <{root.tag}>)")
    base = os.path.dirname(os.path.abspath(path))
    links, joints = [], []
    for le in root.findall("link"):
        inertial = None
        ie = le.find("inertial")
        if ie is not None:
            m = ie.find("mass")
            ii = ie.find("inertia")
            a = ii.attrib if ii is not None else {}
            g = lambda k: float(a.get(k, 0.0))  # noqa: E731
            I = np.array([[g("ixx"), g("ixy"), g("ixz")], [g("ixy"), g("iyy"), g("iyz")], [g("ixz"), g("iyz"), g("izz")]])
            inertial = _Inertial(float(m.attrib["value"]) if m is not None else 0.0, _origin(ie), I)
        link = _Link(le.attrib["name"], inertial)
        for ce in le.findall("collision"):
            ge = ce.find("geometry")
            shape = list(ge)[0] if ge is not None and len(ge) else None
            if shape is None:
                continue
            if shape.tag == "box":
                col = _Collision(_origin(ce), "box", _vec(shape.attrib["size"]))
            elif shape.tag == "sphere":
                col = _Collision(_origin(ce), "sphere", np.array([float(shape.attrib["radius"])]))
            elif shape.tag == "mesh":
                fn = shape.attrib["filename"]
                if fn.startswith("package://"):
                    fn = fn[len("package://"):]
                if not os.path.isabs(fn):
                    fn = os.path.join(base, fn)
                col = _Collision(_origin(ce), "mesh", _vec(shape.attrib.get("scale"), 3, 1.0), fn)
            else:
                raise NotImplementedError(f"URDF collision geometry <{shape.tag}> is not supported")
            link.collisions.append(col)
        links.append(link)
    for je in root.findall("joint"):
        jt = je.attrib["type"]
        if jt not in ("revolute", "continuous", "prismatic", "fixed", "floating"):
            raise NotImplementedError(f"URDF joint type '{jt}' is not supported")
        ax = je.find("axis")
        axis = _vec(ax.attrib.get("xyz"), 3) if ax is not None else np.array([1.0, 0.0, 0.0])
        j = _Joint(je.attrib["name"], jt, je.find("parent").attrib["link"], je.find("child").attrib["link"],
                   _origin(je), axis)
        lim = je.find("limit")
        if lim is not None:
            j.lower = float(lim.attrib["lower"]) if "lower" in lim.attrib else None
            j.upper = float(lim.attrib["upper"]) if "upper" in lim.attrib else None
            j.effort = float(lim.attrib["effort"]) if "effort" in lim.attrib else None
        dyn = je.find("dynamics")
        if dyn is not None:
            j.damping = float(dyn.attrib.get("damping", 0.0))
            j.friction = float(dyn.attrib.get("friction", 0.0))
        sc = je.find("safety_controller")
        if sc is not None:
            j.k_position = float(sc.attrib["k_position"]) if "k_position" in sc.attrib else None
            j.k_velocity = float(sc.attrib["k_velocity"]) if "k_velocity" in sc.attrib else None
        joints.append(j)
    return root.attrib.get("name", "robot"), links, joints


def _base_link(links, joints):
    children = {j.child for j in joints}
    roots = [l.name for l in links if l.name not in children]
    if len(roots) != 1:
        raise ValueError(f"URDF: expected one root link, found {roots}")
    return roots[0]


def _merge_inertia(parent: _Link, child: _Link):
    """urdf.py:572-611: the child's inertial (already in the parent's frame) folded into the parent's, about the
    combined centre of mass, in the parent frame's axes."""
    if child.inertial is None:
        return
    if parent.inertial is None:
        parent.inertial = child.inertial
        return
    m1, m2 = parent.inertial.mass, child.inertial.mass
    c1, R1 = parent.inertial.origin[:3, 3], parent.inertial.origin[:3, :3]
    c2, R2 = child.inertial.origin[:3, 3], child.inertial.origin[:3, :3]
    m = m1 + m2
    com = (m1 * c1 + m2 * c2) / m if m > np.finfo(np.float32).eps else c1

    def moved(I, mass, d):
        return I + mass * (np.dot(d, d) * np.eye(3) - np.outer(d, d))

    I = moved(R1 @ parent.inertial.inertia @ R1.T, m1, com - c1) + moved(R2 @ child.inertial.inertia @ R2.T, m2, com - c2)
    origin = np.eye(4)
    origin[:3, 3] = com
    parent.inertial = _Inertial(m, origin, I)


def merge_fixed_links(links, joints, links_to_keep=()):
    """urdf.py:436-472: the fixed joints, in reverse breadth-first order from the base, each fold its child link
    into its parent (inertial, collision geometry, and the child's own joints re-expressed in the parent's frame)."""
    links = [copy.deepcopy(l) for l in links]
    joints = [copy.deepcopy(j) for j in joints]
    by_parent = {}
    for j in joints:
        by_parent.setdefault(j.parent, []).append(j)
    ordered, queue, seen = [], [_base_link(links, joints)], set()
    while queue:
        name = queue.pop(0)
        if name in seen:
            continue
        seen.add(name)
        for j in by_parent.get(name, []):
            ordered.append(j)
            queue.append(j.child)
    for j in ordered[::-1]:
        if j.type != "fixed" or j.child in links_to_keep:
            continue
        parent = next(l for l in links if l.name == j.parent)
        child = next(l for l in links if l.name == j.child)
        T = j.origin
        if child.inertial is not None:
            child.inertial = _Inertial(child.inertial.mass, T @ child.inertial.origin, child.inertial.inertia)
        for c in child.collisions:
            c.origin = T @ c.origin
        for jj in joints:
            if jj.parent == child.name:
                jj.origin = T @ jj.origin
        _merge_inertia(parent, child)
        parent.collisions.extend(child.collisions)
        joints.remove(j)
        links.remove(child)
        for jj in joints:
            if jj.parent == child.name:
                jj.parent = parent.name
    return links, joints


def _fmt(v):
    return " ".join(repr(float(x)) for x in np.ravel(v))


def to_mjcf(name, links, joints, fixed: bool, pos=(0.0, 0.0, 0.0), quat=(1.0, 0.0, 0.0, 0.0)) -> ET.Element:
    """The ``<mujoco>`` tree MuJoCo compiles from the (merged) URDF under the reference's compiler flags."""
    mj = ET.Element("mujoco", model=name)
    ET.SubElement(mj, "compiler", angle="radian", autolimits="true")
    asset = ET.SubElement(mj, "asset")
    world = ET.SubElement(mj, "worldbody")
    by_name = {l.name: l for l in links}
    by_parent = {}
    for j in joints:
        by_parent.setdefault(j.parent, []).append(j)
    n_mesh = [0]

    def body(elem, link, T, joint):
        b = ET.SubElement(elem, "body", name=link.name, pos=_fmt(T[:3, 3]), quat=_fmt(tf.R_to_quat(T[:3, :3])))
        if link.inertial is not None:
            o = link.inertial.origin
            I = link.inertial.inertia
            ET.SubElement(b, "inertial", pos=_fmt(o[:3, 3]), quat=_fmt(tf.R_to_quat(o[:3, :3])),
                          mass=repr(float(link.inertial.mass)),
                          fullinertia=_fmt([I[0, 0], I[1, 1], I[2, 2], I[0, 1], I[0, 2], I[1, 2]]))
        else:  # inertiafromgeom="false": no inertial means no mass (build_model zeroes MuJoCo's placeholder)
            ET.SubElement(b, "inertial", pos="0 0 0", mass="0", diaginertia="0 0 0")
        if joint is not None and joint.type != "fixed":
            if joint.type == "floating":
                ET.SubElement(b, "freejoint", name=joint.name)
            else:
                attrs = {"name": joint.name, "damping": repr(joint.damping), "frictionloss": repr(joint.friction),
                         "type": "slide" if joint.type == "prismatic" else "hinge",
                         "axis": _fmt(joint.axis / np.linalg.norm(joint.axis))}
                if joint.type != "continuous" and joint.lower is not None and joint.upper is not None:
                    attrs["range"] = _fmt([joint.lower, joint.upper])
                ET.SubElement(b, "joint", **attrs)
        elif joint is None and not fixed:
            ET.SubElement(b, "freejoint", name="root_joint")
        for c in link.collisions:
            g = {"pos": _fmt(c.origin[:3, 3]), "quat": _fmt(tf.R_to_quat(c.origin[:3, :3]))}
            if c.kind == "box":
                g.update(type="box", size=_fmt(0.5 * c.size))
            elif c.kind == "sphere":
                g.update(type="sphere", size=_fmt(c.size))
            else:
                mname = f"mesh{n_mesh[0]}"
                n_mesh[0] += 1
                ET.SubElement(asset, "mesh", name=mname, file=c.filename, scale=_fmt(c.size))
                g.update(type="mesh", mesh=mname)
            ET.SubElement(b, "geom", **g)
        for jj in by_parent.get(link.name, []):
            body(b, by_name[jj.child], jj.origin, jj)

    T0 = np.eye(4)
    T0[:3, :3] = tf.quat_to_R(np.asarray(quat, dtype=np.float64))
    T0[:3, 3] = pos
    body(world, by_name[_base_link(links, joints)], T0, None)
    return mj


def load_urdf(path: str, name: str | None = None, fixed: bool = False, merge_fixed: bool = True,
              links_to_keep=(), default_armature: float | None = DEFAULT_ARMATURE, pos=(0.0, 0.0, 0.0),
              quat=(1.0, 0.0, 0.0, 0.0)) -> EntitySpec:
    """Compile a URDF file into an entity, with ``gs.morphs.URDF``'s options (``fixed``, ``merge_fixed_links``,
    ``links_to_keep``, ``default_armature``, ``pos``/``quat`` of the base)."""
    robot, links, joints = _parse(path)
    if merge_fixed:
        links, joints = merge_fixed_links(links, joints, tuple(links_to_keep))
    tree = to_mjcf(robot, links, joints, fixed, pos, quat)
    spec = MJCFLoader(path, root=tree).load(name or robot)
    # what the reference's own URDF parser supplies over MuJoCo's (rigid_entity.py:540-548, urdf.py:366-387)
    by_name = {j.name: j for j in joints}
    for link in spec.links:
        for js in link.joints:
            uj = by_name.get(js.name)
            if uj is None:  # the free root joint the morph adds: no gains, no armature (rigid_entity.py:662-686)
                continue
            if js.type in (JOINT_REVOLUTE, JOINT_PRISMATIC):
                js.armature = float(default_armature) if default_armature is not None else 0.0
            elif js.type == JOINT_FREE:
                js.armature = 0.0
            kp = uj.k_position if uj.k_position is not None else DEFAULT_KP
            kv = uj.k_velocity if uj.k_velocity is not None else DEFAULT_KV
            js.act_gain = kp
            js.act_bias = (0.0, -kp, -kv)
            js.force_range = (-uj.effort, uj.effort) if uj.effort is not None else (-np.inf, np.inf)
    return spec
