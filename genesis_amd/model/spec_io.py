"""Save / load an EntitySpec (the loader's output: links, joints, collision geoms with convex-hull vertices).

A saved spec lets a model be rebuilt where its MJCF source tree is not available (the GPU host only receives this
repository): the compiler then regenerates every derived table (inertia weights, pair filter, support tables).
"""

from __future__ import annotations

import json

import numpy as np

from .specs import EntitySpec, EqualitySpec, GeomSpec, JointSpec, LinkSpec

FORMAT_VERSION = 1


def save_spec(spec: EntitySpec, path: str) -> None:
    arrays = {}
    meta = {"version": FORMAT_VERSION, "name": spec.name, "is_local_collision_mask": spec.is_local_collision_mask,
            "gravity_compensation": spec.gravity_compensation, "links": [], "equalities": []}
    for i, l in enumerate(spec.links):
        lm = {"name": l.name, "parent": l.parent, "inertial_mass": float(l.inertial_mass), "joints": [], "geoms": []}
        for key in ("pos", "quat", "inertial_pos", "inertial_quat", "inertial_i"):
            arrays[f"l{i}_{key}"] = np.asarray(getattr(l, key), np.float64)
        for jn, j in enumerate(l.joints):
            jm = {"name": j.name, "type": j.type, "limit": None if j.limit is None else [float(x) for x in j.limit],
                  "armature": j.armature, "damping": j.damping, "stiffness": j.stiffness,
                  "frictionloss": j.frictionloss, "act_gain": float(j.act_gain),
                  "act_bias": [float(x) for x in j.act_bias], "force_range": [float(x) for x in j.force_range]}
            arrays[f"l{i}_j{jn}_pos"] = np.asarray(j.pos, np.float64)
            arrays[f"l{i}_j{jn}_axis"] = np.asarray(j.axis, np.float64)
            arrays[f"l{i}_j{jn}_sol_params"] = np.asarray(j.sol_params, np.float64)
            if j.init_qpos is not None:
                arrays[f"l{i}_j{jn}_init_qpos"] = np.asarray(j.init_qpos, np.float64)
            lm["joints"].append(jm)
        for gn, g in enumerate(l.geoms):
            gm = {"type": g.type, "friction": g.friction, "contype": g.contype, "conaffinity": g.conaffinity,
                  "name": g.name, "has_faces": g.faces is not None, "has_center": g.center is not None}
            for key in ("pos", "quat", "data", "verts", "sol_params"):
                arrays[f"l{i}_g{gn}_{key}"] = np.asarray(getattr(g, key), np.float64)
            if g.faces is not None:
                arrays[f"l{i}_g{gn}_faces"] = np.asarray(g.faces, np.int64)
            if g.center is not None:
                arrays[f"l{i}_g{gn}_center"] = np.asarray(g.center, np.float64)
            lm["geoms"].append(gm)
        meta["links"].append(lm)
    for k, e in enumerate(spec.equalities):
        meta["equalities"].append({"type": e.type, "obj1": e.obj1, "obj2": e.obj2})
        arrays[f"e{k}_data"] = np.asarray(e.data, np.float64)
        arrays[f"e{k}_sol_params"] = np.asarray(e.sol_params, np.float64)
    arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez_compressed(path, **arrays)


def load_spec(path: str) -> EntitySpec:
    z = np.load(path, allow_pickle=False)
    meta = json.loads(bytes(z["meta"]).decode())
    if meta["version"] != FORMAT_VERSION:
        raise ValueError(f"{path}: spec format {meta['version']} != {FORMAT_VERSION}")
    links = []
    for i, lm in enumerate(meta["links"]):
        l = LinkSpec(name=lm["name"], parent=lm["parent"], pos=z[f"l{i}_pos"], quat=z[f"l{i}_quat"],
                     inertial_pos=z[f"l{i}_inertial_pos"], inertial_quat=z[f"l{i}_inertial_quat"],
                     inertial_i=z[f"l{i}_inertial_i"], inertial_mass=lm["inertial_mass"])
        for jn, jm in enumerate(lm["joints"]):
            l.joints.append(JointSpec(
                name=jm["name"], type=jm["type"], pos=z[f"l{i}_j{jn}_pos"], axis=z[f"l{i}_j{jn}_axis"],
                limit=None if jm["limit"] is None else tuple(jm["limit"]), armature=jm["armature"],
                damping=jm["damping"], stiffness=jm["stiffness"], frictionloss=jm["frictionloss"],
                sol_params=z[f"l{i}_j{jn}_sol_params"], act_gain=jm["act_gain"], act_bias=tuple(jm["act_bias"]),
                force_range=tuple(jm["force_range"]),
                init_qpos=z[f"l{i}_j{jn}_init_qpos"] if f"l{i}_j{jn}_init_qpos" in z else None))
        for gn, gm in enumerate(lm["geoms"]):
            l.geoms.append(GeomSpec(
                type=gm["type"], pos=z[f"l{i}_g{gn}_pos"], quat=z[f"l{i}_g{gn}_quat"], data=z[f"l{i}_g{gn}_data"],
                verts=z[f"l{i}_g{gn}_verts"], faces=z[f"l{i}_g{gn}_faces"] if gm["has_faces"] else None,
                friction=gm["friction"], sol_params=z[f"l{i}_g{gn}_sol_params"], contype=gm["contype"],
                conaffinity=gm["conaffinity"], center=z[f"l{i}_g{gn}_center"] if gm["has_center"] else None,
                name=gm["name"]))
        links.append(l)
    eqs = [EqualitySpec(type=em["type"], obj1=em["obj1"], obj2=em["obj2"], data=z[f"e{k}_data"],
                        sol_params=z[f"e{k}_sol_params"]) for k, em in enumerate(meta["equalities"])]
    return EntitySpec(name=meta["name"], links=links, equalities=eqs,
                      is_local_collision_mask=meta["is_local_collision_mask"],
                      gravity_compensation=meta["gravity_compensation"])
