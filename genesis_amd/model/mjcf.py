"""MJCF subset loader.

The reference hands MJCF files to MuJoCo's compiler and then reads the compiled ``mjModel``
(``genesis/utils/mjcf.py:62-198`` ``build_model``, ``:229-404`` ``parse_link``, ``:441-615`` ``parse_geom``).
MuJoCo is not part of this build, so this module restates the parts of that compilation the rigid step depends on:
default-class inheritance, inertial principal axes, joint/actuator parameters (affine-bias ``general`` actuators map to
``act_gain``/``act_bias``; a fixed tendon driven by one actuator is applied to each of its joints, ``mjcf.py:333-349``),
joint equalities, and MuJoCo's convention of expressing mesh geometry in the mesh's own inertial frame (centre of mass
and principal axes). Visual-only geoms (``contype == conaffinity == 0``) are dropped. Links are re-ordered breadth-first
(``genesis/utils/urdf.py:90-128`` ``_order_links``) and the implicit world body is removed when it carries nothing
(``rigid_entity.py:622-639``). Empty joints are dropped (``rigid_entity.py:776``).
"""

from __future__ import annotations

import copy
import os
import xml.etree.ElementTree as ET

import numpy as np

from . import meshes as mu
from . import transforms as tf
from .specs import (
    DEFAULT_SOLIMP,
    EQ_CONNECT,
    EQ_JOINT,
    EQ_WELD,
    GEOM_BOX,
    GEOM_CAPSULE,
    GEOM_CYLINDER,
    GEOM_ELLIPSOID,
    GEOM_MESH,
    GEOM_PLANE,
    GEOM_SPHERE,
    JOINT_FREE,
    JOINT_PRISMATIC,
    JOINT_REVOLUTE,
    JOINT_SPHERICAL,
    MIN_TIMECONST,
    EntitySpec,
    EqualitySpec,
    GeomSpec,
    JointSpec,
    LinkSpec,
)

_GEOM_TYPES = {
    "plane": GEOM_PLANE,
    "sphere": GEOM_SPHERE,
    "ellipsoid": GEOM_ELLIPSOID,
    "cylinder": GEOM_CYLINDER,
    "capsule": GEOM_CAPSULE,
    "box": GEOM_BOX,
    "mesh": GEOM_MESH,
}
_JOINT_TYPES = {"hinge": JOINT_REVOLUTE, "slide": JOINT_PRISMATIC, "ball": JOINT_SPHERICAL, "free": JOINT_FREE}


def _floats(s, n=None):
    v = np.array([float(x) for x in s.split()], dtype=np.float64)
    if n is not None and len(v) < n:
        v = np.concatenate([v, np.zeros(n - len(v))])
    return v


class _Defaults:
    """One MuJoCo default class: attribute dictionaries per element tag."""

    def __init__(self, parent=None):
        self.attrs = {} if parent is None else copy.deepcopy(parent.attrs)

    def update(self, elem):
        for child in elem:
            if child.tag == "default":
                continue
            self.attrs.setdefault(child.tag, {}).update(child.attrib)


class MJCFLoader:
    def __init__(self, path: str, root: ET.Element | None = None):
        """``root``: an already-built ``<mujoco>`` tree (the URDF loader's translation); relative mesh files are then
        resolved against ``path``'s directory as for a file."""
        self.path = path
        self.root = ET.parse(path).getroot() if root is None else root
        comp = self.root.find("compiler")
        self.angle_deg = comp is not None and comp.attrib.get("angle", "degree") == "degree"
        self.autolimits = comp is None or comp.attrib.get("autolimits", "true") == "true"
        meshdir = comp.attrib.get("meshdir", "") if comp is not None else ""
        self.meshdir = os.path.join(os.path.dirname(path), meshdir)
        self.classes = {}
        main = _Defaults()
        # The reference injects near-zero time constants into the top-level defaults so that the solver's own
        # default time constant is used unless the file overrides it (genesis/utils/mjcf.py:118-131).
        main.attrs["geom"] = {"solref": f"{MIN_TIMECONST} 1"}
        main.attrs["joint"] = {"solreflimit": f"{MIN_TIMECONST} 1", "solreffriction": f"{MIN_TIMECONST} 1"}
        main.attrs["equality"] = {"solref": f"{MIN_TIMECONST} 1"}
        top = self.root.find("default")
        if top is not None:
            main.update(top)
            self.classes["main"] = main
            self._collect(top, main)
        else:
            self.classes["main"] = main
        self.mesh_files = {}
        self.mesh_scale = {}
        asset = self.root.find("asset")
        if asset is not None:
            for m in asset.findall("mesh"):
                f = m.attrib["file"]
                name = m.attrib.get("name", os.path.splitext(os.path.basename(f))[0])
                self.mesh_files[name] = os.path.join(self.meshdir, f)
                if "scale" in m.attrib:
                    self.mesh_scale[name] = _floats(m.attrib["scale"], 3)
        self._mesh_cache = {}

    def _collect(self, elem, parent):
        for child in elem.findall("default"):
            d = _Defaults(parent)
            d.update(child)
            self.classes[child.attrib.get("class", "main")] = d
            self._collect(child, d)

    def _attr(self, elem, tag, cls, key, default=None):
        if key in elem.attrib:
            return elem.attrib[key]
        c = self.classes.get(elem.attrib.get("class", cls), self.classes["main"])
        return c.attrs.get(tag, {}).get(key, default)

    def _angle(self, v):
        return np.deg2rad(v) if self.angle_deg else v

    def _orientation(self, elem, tag, cls):
        q = self._attr(elem, tag, cls, "quat")
        if q is not None:
            return tf.quat_normalize(_floats(q))
        e = self._attr(elem, tag, cls, "euler")
        if e is not None:
            return tf.euler_xyz_to_quat(self._angle(_floats(e)))
        aa = self._attr(elem, tag, cls, "axisangle")
        if aa is not None:
            v = _floats(aa)
            return tf.axis_angle_quat(v[:3], self._angle(v[3]))
        return np.array([1.0, 0.0, 0.0, 0.0])

    # ------------------------------------------------------------------ meshes
    def _mesh(self, name):
        if name not in self._mesh_cache:
            verts, faces = mu.load_mesh(self.mesh_files[name])
            if name in self.mesh_scale:
                verts = verts * self.mesh_scale[name]
            # MuJoCo stores mesh vertices in the mesh's inertial frame; the geom pose absorbs that frame.
            vol, com, inertia = mu.mass_properties(verts, faces)
            if vol <= 0.0:
                hv, hf = mu.convex_hull(verts)
                vol, com, inertia = mu.mass_properties(hv, hf)
            R = mu.principal_frame(inertia)
            local = (verts - com) @ R
            hull_v, hull_f = mu.convex_hull(local)
            self._mesh_cache[name] = (com, tf.R_to_quat(R), hull_v, hull_f)
        return self._mesh_cache[name]

    # ------------------------------------------------------------------ bodies
    def load(self, name="entity") -> EntitySpec:
        world = self.root.find("worldbody")
        bodies = []  # (elem, parent_index, childclass)

        def walk(elem, parent_idx, cls):
            for b in elem.findall("body"):
                bcls = b.attrib.get("childclass", cls)
                idx = len(bodies)
                bodies.append((b, parent_idx, bcls))
                walk(b, idx, bcls)

        world_geoms = [g for g in world.findall("geom")]
        walk(world, -1, "main")
        links = []
        joint_owner = {}
        for b, parent, cls in bodies:
            links.append(self._body(b, parent, cls, joint_owner))
        if world_geoms:
            raise NotImplementedError("geoms attached to the MJCF world body are not supported")
        links = self._order(links)
        equalities = self._equalities()
        self._actuators(links)
        for link in links:
            link.joints = [j for j in link.joints if j.n_dofs > 0]
        return EntitySpec(name=name, links=links, equalities=equalities, is_local_collision_mask=True)

    def _body(self, b, parent, cls, joint_owner):
        pos = _floats(b.attrib.get("pos", "0 0 0"))
        quat = self._orientation(b, "body", cls)
        inertial = b.find("inertial")
        if inertial is None:
            mass, ipos, iquat, inertia = self._inertia_from_geoms(b, cls)
        else:
            mass = float(inertial.attrib["mass"])
            ipos = _floats(inertial.attrib.get("pos", "0 0 0"))
            iquat = self._orientation(inertial, "inertial", cls)
            if "fullinertia" in inertial.attrib:
                ixx, iyy, izz, ixy, ixz, iyz = _floats(inertial.attrib["fullinertia"])
                I = np.array([[ixx, ixy, ixz], [ixy, iyy, iyz], [ixz, iyz, izz]])
                I = tf.quat_to_R(iquat) @ I @ tf.quat_to_R(iquat).T
                evals, evecs = np.linalg.eigh(I)
                if np.linalg.det(evecs) < 0:
                    evecs[:, 2] = -evecs[:, 2]
                iquat = tf.R_to_quat(evecs)
                inertia = np.diag(evals)
            else:
                inertia = np.diag(_floats(inertial.attrib["diaginertia"]))
        link = LinkSpec(
            name=b.attrib.get("name", ""),
            parent=parent,
            pos=pos,
            quat=quat,
            inertial_pos=ipos,
            inertial_quat=iquat,
            inertial_i=inertia,
            inertial_mass=mass,
        )
        for j in b:
            if j.tag == "freejoint":
                link.joints.append(JointSpec(name=j.attrib.get("name", link.name), type=JOINT_FREE))
            elif j.tag == "joint":
                link.joints.append(self._joint(j, cls))
        for g in b.findall("geom"):
            spec = self._geom(g, cls)
            if spec is not None:
                link.geoms.append(spec)
        return link

    def _fromto(self, g, cls, size, pos, quat):
        """A geom given by `fromto`: centred on the segment, its z axis along it, half-length size[1] (set in
        place); else (pos, quat) unchanged."""
        fromto = self._attr(g, "geom", cls, "fromto")
        if fromto is None:
            return pos, quat
        ft = _floats(fromto)
        a, b = ft[:3], ft[3:]
        d = b - a
        size[1] = 0.5 * np.linalg.norm(d)
        z = np.array([0.0, 0.0, 1.0])
        dn = d / np.linalg.norm(d)
        c = np.cross(z, dn)
        if np.linalg.norm(c) < 1e-12:
            q = np.array([1.0, 0.0, 0.0, 0.0]) if dn[2] > 0 else np.array([0.0, 1.0, 0.0, 0.0])
        else:
            q = tf.axis_angle_quat(c, np.arctan2(np.linalg.norm(c), np.dot(z, dn)))
        return 0.5 * (a + b), q

    def _inertia_from_geoms(self, b, cls):
        """MuJoCo's inertiafromgeom: every geom of the body contributes density x volume (or its mass attribute)
        and its inertia about the body frame; the sum is diagonalised into a principal inertial frame."""
        total, first, second = 0.0, np.zeros(3), np.zeros((3, 3))
        parts = []
        for g in b.findall("geom"):
            tname = self._attr(g, "geom", cls, "type", "sphere")
            size = _floats(self._attr(g, "geom", cls, "size", "0 0 0"), 3)
            pos = _floats(self._attr(g, "geom", cls, "pos", "0 0 0"))
            quat = self._orientation(g, "geom", cls)
            if tname in ("capsule", "cylinder", "box", "ellipsoid"):
                pos, quat = self._fromto(g, cls, size, pos, quat)
            if tname == "box":
                a, bb, c = size
                vol = 8.0 * a * bb * c
                unit = np.diag([bb * bb + c * c, a * a + c * c, a * a + bb * bb]) / 3.0
            elif tname == "sphere":
                r = size[0]
                vol = 4.0 / 3.0 * np.pi * r**3
                unit = np.eye(3) * 0.4 * r * r
            elif tname == "capsule":
                r, L = size[0], 2.0 * size[1]
                v_cyl, v_sph = np.pi * r * r * L, 4.0 / 3.0 * np.pi * r**3
                vol = v_cyl + v_sph
                ixx = (v_cyl * (3 * r * r + L * L) / 12 + v_sph * (0.4 * r * r + 0.25 * L * L + 3 / 8 * r * L)) / vol
                unit = np.diag([ixx, ixx, (0.5 * v_cyl * r * r + 0.4 * v_sph * r * r) / vol])
            elif tname == "mesh":
                com, mquat, verts, faces = self._mesh(self._attr(g, "geom", cls, "mesh"))
                vol, _, inertia = mu.mass_properties(verts, faces)
                pos, quat = tf.compose(pos, quat, com, mquat)
                unit = inertia / max(vol, 1e-18)
            else:
                continue
            m_attr = self._attr(g, "geom", cls, "mass")
            mass = float(m_attr) if m_attr is not None else float(self._attr(g, "geom", cls, "density", "1000")) * vol
            R = tf.quat_to_R(quat)
            parts.append((mass, pos, R @ (mass * unit) @ R.T))
            total += mass
            first += mass * pos
        if total <= 0.0:
            return 0.0, np.zeros(3), np.array([1.0, 0.0, 0.0, 0.0]), np.zeros((3, 3))
        com = first / total
        for mass, pos, I in parts:
            d = pos - com
            second += I + mass * (np.dot(d, d) * np.eye(3) - np.outer(d, d))
        evals, evecs = np.linalg.eigh(second)
        if np.linalg.det(evecs) < 0:
            evecs[:, 2] = -evecs[:, 2]
        return total, com, tf.R_to_quat(evecs), np.diag(evals)

    def _joint(self, j, cls):
        jtype = _JOINT_TYPES[self._attr(j, "joint", cls, "type", "hinge")]
        axis = _floats(self._attr(j, "joint", cls, "axis", "0 0 1"))
        axis = axis / np.linalg.norm(axis)
        rng = self._attr(j, "joint", cls, "range")
        limited = self._attr(j, "joint", cls, "limited")
        is_limited = (limited == "true") or (limited is None and self.autolimits and rng is not None)
        limit = None
        if is_limited and rng is not None:
            lo, hi = _floats(rng)
            if jtype == JOINT_REVOLUTE:
                lo, hi = self._angle(lo), self._angle(hi)
            limit = (lo, hi)
        solref = _floats(self._attr(j, "joint", cls, "solreflimit", f"{MIN_TIMECONST} 1"))
        solimp = _floats(self._attr(j, "joint", cls, "solimplimit", " ".join(map(str, DEFAULT_SOLIMP))))
        solimp = np.concatenate([solimp, np.array(DEFAULT_SOLIMP)[len(solimp) :]])
        return JointSpec(
            name=j.attrib.get("name", ""),
            type=jtype,
            pos=_floats(self._attr(j, "joint", cls, "pos", "0 0 0")),
            axis=axis,
            limit=limit,
            armature=float(self._attr(j, "joint", cls, "armature", "0")),
            damping=float(self._attr(j, "joint", cls, "damping", "0")),
            stiffness=float(self._attr(j, "joint", cls, "stiffness", "0")),
            frictionloss=float(self._attr(j, "joint", cls, "frictionloss", "0")),
            sol_params=np.concatenate([solref[:2], solimp[:5]]),
        )

    def _geom(self, g, cls):
        contype = int(self._attr(g, "geom", cls, "contype", "1"))
        conaffinity = int(self._attr(g, "geom", cls, "conaffinity", "1"))
        if not (contype or conaffinity):
            return None
        tname = self._attr(g, "geom", cls, "type", "sphere")
        gtype = _GEOM_TYPES[tname]
        size = _floats(self._attr(g, "geom", cls, "size", "0 0 0"), 3)
        pos = _floats(self._attr(g, "geom", cls, "pos", "0 0 0"))
        quat = self._orientation(g, "geom", cls)
        if gtype in (GEOM_CAPSULE, GEOM_CYLINDER, GEOM_BOX, GEOM_ELLIPSOID):
            pos, quat = self._fromto(g, cls, size, pos, quat)
        friction = float(_floats(self._attr(g, "geom", cls, "friction", "1 0.005 0.0001"))[0])
        solref = _floats(self._attr(g, "geom", cls, "solref", f"{MIN_TIMECONST} 1"))
        solimp = _floats(self._attr(g, "geom", cls, "solimp", " ".join(map(str, DEFAULT_SOLIMP))))
        solimp = np.concatenate([solimp, np.array(DEFAULT_SOLIMP)[len(solimp) :]])
        sol_params = np.concatenate([solref[:2], solimp[:5]])
        faces = None
        if gtype == GEOM_MESH:
            com, mquat, verts, faces = self._mesh(self._attr(g, "geom", cls, "mesh"))
            pos, quat = tf.compose(pos, quat, com, mquat)
            data = np.zeros(7)
        elif gtype == GEOM_BOX:
            data = 2.0 * size
            verts, faces = mu.box_mesh(data)
        elif gtype == GEOM_SPHERE:
            data = np.array([size[0]])
            verts = mu.sphere_points(size[0])
        elif gtype == GEOM_CAPSULE:
            data = np.array([size[0], 2.0 * size[1]])
            verts = mu.capsule_points(size[0], 2.0 * size[1])
        elif gtype == GEOM_PLANE:
            raise NotImplementedError("MJCF plane geoms: use the Plane morph")
        else:
            raise NotImplementedError(f"MJCF geom type '{tname}' is not supported")
        return GeomSpec(
            type=gtype,
            pos=pos,
            quat=quat,
            data=np.asarray(data, dtype=np.float64),
            verts=verts,
            faces=faces,
            friction=friction,
            sol_params=sol_params,
            contype=contype,
            conaffinity=conaffinity,
            name=g.attrib.get("name", ""),
        )

    # ---------------------------------------------------------------- ordering
    @staticmethod
    def _order(links):
        children = {i: [] for i in range(len(links))}
        for i, l in enumerate(links):
            if l.parent >= 0:
                children[l.parent].append(i)
        order, level = [], [i for i, l in enumerate(links) if l.parent < 0]
        while level:
            order += level
            level = [c for i in level for c in children[i]]
        new_index = {old: new for new, old in enumerate(order)}
        out = []
        for old in order:
            l = links[old]
            l.parent = new_index[l.parent] if l.parent >= 0 else -1
            out.append(l)
        return out

    # -------------------------------------------------------------- equality
    def _equalities(self):
        out = []
        eq = self.root.find("equality")
        if eq is None:
            return out
        for e in eq:
            solref = _floats(e.attrib.get("solref", f"{MIN_TIMECONST} 1"))
            solimp = _floats(e.attrib.get("solimp", " ".join(map(str, DEFAULT_SOLIMP))))
            solimp = np.concatenate([solimp, np.array(DEFAULT_SOLIMP)[len(solimp) :]])
            sol_params = np.concatenate([solref[:2], solimp[:5]])
            data = np.zeros(11)
            if e.tag == "joint":
                poly = _floats(e.attrib.get("polycoef", "0 1 0 0 0"), 5)
                data[:5] = poly[:5]
                out.append(EqualitySpec(type=EQ_JOINT, obj1=e.attrib["joint1"], obj2=e.attrib["joint2"], data=data,
                                        sol_params=sol_params))
            elif e.tag in ("connect", "weld"):
                # body-to-body only (mjcf.py parse_equalities takes the compiled MuJoCo eq_data: the second anchor,
                # and a weld's unset relative pose, follow from qpos0)
                if "body1" not in e.attrib or "body2" not in e.attrib:
                    raise NotImplementedError(f"MJCF {e.tag}: only body1/body2 equalities are supported")
                data[0:3] = _floats(e.attrib.get("anchor", "0 0 0"), 3)
                if e.tag == "connect":
                    etype = EQ_CONNECT
                else:
                    etype = EQ_WELD
                    relpose = _floats(e.attrib.get("relpose", "0 1 0 0 0 0 0"), 7)
                    if np.any(relpose[3:7]) and not np.allclose(relpose, [0, 1, 0, 0, 0, 0, 0]):
                        raise NotImplementedError("MJCF weld: an explicit relpose is not supported")
                    data[10] = float(_floats(e.attrib.get("torquescale", "1"), 1)[0])
                out.append(EqualitySpec(type=etype, obj1=e.attrib["body1"], obj2=e.attrib["body2"], data=data,
                                        sol_params=sol_params, at_qpos0=True))
            else:
                raise NotImplementedError(f"MJCF equality '{e.tag}' is not supported")
        return out

    # ------------------------------------------------------------- actuators
    def _actuators(self, links):
        joints = {j.name: j for l in links for j in l.joints}
        tendons = {}
        ten = self.root.find("tendon")
        if ten is not None:
            for t in ten.findall("fixed"):
                tendons[t.attrib["name"]] = [jj.attrib["joint"] for jj in t.findall("joint")]
        act = self.root.find("actuator")
        if act is None:
            return
        for a in act:
            cls = a.attrib.get("class", "main")
            tag = a.tag
            gear = float(_floats(self._attr(a, tag, cls, "gear", "1"))[0])
            if tag == "general":
                gainprm = _floats(self._attr(a, tag, cls, "gainprm", "1"), 3)
                biasprm = _floats(self._attr(a, tag, cls, "biasprm", "0"), 3)
                biastype = self._attr(a, tag, cls, "biastype", "none")
                if biastype == "none":
                    biasprm = np.zeros(3)
            elif tag == "position":
                kp = float(self._attr(a, tag, cls, "kp", "1"))
                kv = float(self._attr(a, tag, cls, "kv", "0"))
                gainprm, biasprm, biastype = np.array([kp, 0, 0]), np.array([0.0, -kp, -kv]), "affine"
            elif tag == "motor":
                gainprm, biasprm, biastype = np.array([1.0, 0, 0]), np.zeros(3), "none"
            else:
                raise NotImplementedError(f"MJCF actuator '{tag}' is not supported")
            frange = self._attr(a, tag, cls, "forcerange")
            flimited = self._attr(a, tag, cls, "forcelimited")
            is_flim = (flimited == "true") or (flimited is None and self.autolimits and frange is not None)
            if "joint" in a.attrib:
                targets = [a.attrib["joint"]]
            else:
                targets = tendons[a.attrib["tendon"]]
            for jn in targets:
                j = joints[jn]
                j.act_gain = gear * gainprm[0]
                j.act_bias = tuple(gear * biasprm[:3]) if biastype != "none" else (0.0, 0.0, 0.0)
                if is_flim:
                    j.force_range = tuple(_floats(frange))
                if biastype == "none":
                    crange = self._attr(a, tag, cls, "ctrlrange")
                    climited = self._attr(a, tag, cls, "ctrllimited")
                    if (climited == "true") or (climited is None and self.autolimits and crange is not None):
                        c = gear * _floats(crange)
                        j.force_range = (max(j.force_range[0], c[0]), min(j.force_range[1], c[1]))


def load_mjcf(path: str, name: str | None = None) -> EntitySpec:
    return MJCFLoader(path).load(name or os.path.splitext(os.path.basename(path))[0])
