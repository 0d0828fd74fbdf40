"""Pack a compiled model into the versioned blob of ``include/rigidstep_arrays.h`` (the C ABI's model format).

The array directory is read from that header, so the C++ side and this packer share one definition.
"""

from __future__ import annotations

import ctypes
import os
import re

import numpy as np

from .compiler import CompiledModel

ABI_VERSION = 3
MAX_MODEL_ARRAYS = 96
_HERE = os.path.dirname(os.path.abspath(__file__))
_INCLUDE = os.path.normpath(os.path.join(_HERE, "..", "..", "include"))

COUNT_KINDS = [
    "RS_N_ONE", "RS_N_LINKS", "RS_N_JOINTS", "RS_N_DOFS", "RS_N_QS", "RS_N_GEOMS", "RS_N_ENTITIES",
    "RS_N_EQUALITIES", "RS_N_DOFS2", "RS_N_GEOMS2", "RS_N_PAIRS", "RS_N_SUPPORT", "RS_N_CONTACTS",
    "RS_N_CONSTRAINTS", "RS_N_CONSTRAINTS_DOFS", "RS_N_BROAD", "RS_N_SORT", "RS_N_WORK", "RS_N_SCHED",
    "RS_N_DOFS_PLANE", "RS_N_ONE_PLANE", "RS_N_VERTS",
]


def model_array_directory(path: str | None = None) -> list[tuple[str, bool, str, int]]:
    """[(name, is_int, count_kind, width)] in blob order, parsed from rigidstep_arrays.h."""
    path = path or os.path.join(_INCLUDE, "rigidstep_arrays.h")
    text = open(path).read()
    out = []
    for name, is_int, kind, width in re.findall(r"RS_MA\((\w+),\s*(\d),\s*(RS_N_\w+),\s*(\d+)\)", text):
        out.append((name, bool(int(is_int)), kind, int(width)))
    return out


class ModelHeader(ctypes.Structure):
    """Mirror of ``rs_model_header`` (include/rigidstep.h)."""

    _fields_ = [
        ("abi_version", ctypes.c_int32),
        ("float_bytes", ctypes.c_int32),
        ("n_links", ctypes.c_int32),
        ("n_joints", ctypes.c_int32),
        ("n_dofs", ctypes.c_int32),
        ("n_qs", ctypes.c_int32),
        ("n_geoms", ctypes.c_int32),
        ("n_entities", ctypes.c_int32),
        ("n_equalities", ctypes.c_int32),
        ("n_pairs", ctypes.c_int32),
        ("n_support", ctypes.c_int32),
        ("n_verts", ctypes.c_int32),
        ("max_collision_pairs", ctypes.c_int32),
        ("max_contact_pairs", ctypes.c_int32),
        ("max_collision_pairs_broad", ctypes.c_int32),
        ("n_contacts_per_pair", ctypes.c_int32),
        ("len_constraints", ctypes.c_int32),
        ("dt", ctypes.c_float),
        ("gravity", ctypes.c_float * 3),
        ("iterations", ctypes.c_int32),
        ("tolerance", ctypes.c_float),
        ("ls_iterations", ctypes.c_int32),
        ("ls_tolerance", ctypes.c_float),
        ("enable_collision", ctypes.c_int32),
        ("enable_joint_limit", ctypes.c_int32),
        ("disable_constraint", ctypes.c_int32),
        ("enable_multi_contact", ctypes.c_int32),
        ("box_box_detection", ctypes.c_int32),
        ("integrator", ctypes.c_int32),
        ("solver_type", ctypes.c_int32),
        ("meaninertia", ctypes.c_float),
        ("mc_perturbation", ctypes.c_float),
        ("mc_tolerance", ctypes.c_float),
        ("mpr_to_gjk_overlap_ratio", ctypes.c_float),
        ("n_arrays", ctypes.c_int32),
        ("array_offset", ctypes.c_int64 * MAX_MODEL_ARRAYS),
    ]


def _count(model: CompiledModel, kind: str) -> int:
    return {
        "RS_N_ONE": 1,
        "RS_N_LINKS": model.n_links,
        "RS_N_JOINTS": model.n_joints,
        "RS_N_DOFS": model.n_dofs,
        "RS_N_QS": model.n_qs,
        "RS_N_GEOMS": model.n_geoms,
        "RS_N_ENTITIES": model.n_entities,
        "RS_N_EQUALITIES": model.n_equalities,
        "RS_N_DOFS2": model.n_dofs * model.n_dofs,
        "RS_N_GEOMS2": model.n_geoms * model.n_geoms,
        "RS_N_PAIRS": model.n_pairs,
        "RS_N_SUPPORT": len(model.arrays["support_v"]),
        "RS_N_VERTS": len(model.arrays["verts"]),
    }[kind]


_ALIASES = {"geom_support_start": "support_cell_start"}


def pack_model(model: CompiledModel, float_dtype=np.float32) -> tuple[ModelHeader, bytes]:
    """Header + blob. ``float_dtype`` float32 for the device library, float64 for the FP64/FP32 CPU restatement."""
    float_dtype = np.dtype(float_dtype)
    opt = model.options
    h = ModelHeader()
    h.abi_version = ABI_VERSION
    h.float_bytes = float_dtype.itemsize
    h.n_links, h.n_joints, h.n_dofs, h.n_qs = model.n_links, model.n_joints, model.n_dofs, model.n_qs
    h.n_geoms, h.n_entities, h.n_equalities = model.n_geoms, model.n_entities, model.n_equalities
    h.n_pairs = model.n_pairs
    h.n_support = len(model.arrays["support_v"])
    h.n_verts = len(model.arrays["verts"])
    h.max_collision_pairs = model.max_collision_pairs
    h.max_contact_pairs = model.max_contact_pairs
    h.max_collision_pairs_broad = model.max_collision_pairs_broad
    h.n_contacts_per_pair = model.n_contacts_per_pair
    h.len_constraints = model.len_constraints
    h.dt = opt.dt
    h.gravity[:] = list(opt.gravity)
    h.iterations, h.tolerance = opt.iterations, opt.tolerance
    h.ls_iterations, h.ls_tolerance = opt.ls_iterations, opt.ls_tolerance
    h.enable_collision, h.enable_joint_limit = int(opt.enable_collision), int(opt.enable_joint_limit)
    h.disable_constraint, h.enable_multi_contact = int(opt.disable_constraint), int(opt.enable_multi_contact)
    h.box_box_detection = int(opt.box_box_detection)
    h.integrator, h.solver_type = opt.integrator, opt.constraint_solver
    h.meaninertia = model.meaninertia
    h.mc_perturbation, h.mc_tolerance, h.mpr_to_gjk_overlap_ratio = 1e-2, 1e-2, 0.25
    directory = model_array_directory()
    h.n_arrays = len(directory)
    chunks = []
    off = 0
    for i, (name, is_int, kind, width) in enumerate(directory):
        arr = np.asarray(model.arrays[_ALIASES.get(name, name)])
        n = _count(model, kind) * width
        dt = np.int32 if is_int else float_dtype
        data = np.ascontiguousarray(arr, dtype=dt).reshape(-1)
        if data.size != n:
            raise ValueError(f"model array {name}: {data.size} elements, expected {n}")
        pad = (-off) % 64
        chunks.append(b"\0" * pad)
        off += pad
        h.array_offset[i] = off
        b = data.tobytes()
        chunks.append(b)
        off += len(b)
    return h, b"".join(chunks)
