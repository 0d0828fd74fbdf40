"""Mesh loading and mass properties for the model compiler.

Collision meshes are reduced to their convex hull, as the reference does for MJCF robots with the default
``convexify=True`` (``genesis/utils/mesh.py:315-520`` ``postprocess_collision_geoms``; the hull itself comes from
``trimesh.convex.convex_hull``, i.e. qhull, which ``scipy.spatial.ConvexHull`` also wraps).
"""

from __future__ import annotations

import struct

import numpy as np
from scipy.spatial import ConvexHull


def load_stl(path: str) -> tuple[np.ndarray, np.ndarray]:
    """Return (vertices[N,3] float64, faces[M,3] int) of a binary or ASCII STL, with duplicate vertices merged."""
    with open(path, "rb") as f:
        data = f.read()
    tris = None
    if len(data) >= 84:
        n = struct.unpack("<I", data[80:84])[0]
        if 84 + 50 * n == len(data):
            rec = np.frombuffer(
                data[84:], dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]), count=n
            )
            tris = rec["v"].astype(np.float64)
    if tris is None:
        pts = []
        for line in data.decode("ascii", errors="ignore").splitlines():
            line = line.strip()
            if line.startswith("vertex"):
                pts.append([float(x) for x in line.split()[1:4]])
        tris = np.asarray(pts, dtype=np.float64).reshape(-1, 3, 3)
    flat = tris.reshape(-1, 3)
    verts, inv = np.unique(flat, axis=0, return_inverse=True)
    return verts, inv.reshape(-1, 3)


def load_obj(path: str) -> tuple[np.ndarray, np.ndarray]:
    """Return (vertices, triangle faces) of a Wavefront OBJ (polygons fan-triangulated)."""
    verts, faces = [], []
    with open(path) as f:
        for line in f:
            if line.startswith("v "):
                verts.append([float(x) for x in line.split()[1:4]])
            elif line.startswith("f "):
                idx = [int(tok.split("/")[0]) for tok in line.split()[1:]]
                idx = [i - 1 if i > 0 else len(verts) + i for i in idx]
                for k in range(1, len(idx) - 1):
                    faces.append([idx[0], idx[k], idx[k + 1]])
    return np.asarray(verts, dtype=np.float64), np.asarray(faces, dtype=np.int64).reshape(-1, 3)


def load_mesh(path: str) -> tuple[np.ndarray, np.ndarray]:
    if path.lower().endswith(".stl"):
        return load_stl(path)
    if path.lower().endswith(".obj"):
        return load_obj(path)
    raise ValueError(f"unsupported mesh format: {path}")


def convex_hull(verts: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Convex hull as (hull vertices, outward-oriented triangle faces indexing those vertices)."""
    hull = ConvexHull(verts)
    used = np.unique(hull.simplices)
    remap = -np.ones(len(verts), dtype=np.int64)
    remap[used] = np.arange(len(used))
    hv = verts[used]
    faces = remap[hull.simplices]
    centroid = hv.mean(axis=0)
    for k, (a, b, c) in enumerate(faces):
        n = np.cross(hv[b] - hv[a], hv[c] - hv[a])
        if np.dot(n, hv[a] - centroid) < 0.0:
            faces[k] = (a, c, b)
    return hv, faces


def mass_properties(verts: np.ndarray, faces: np.ndarray) -> tuple[float, np.ndarray, np.ndarray]:
    """Volume, centre of mass and inertia tensor (about the COM, unit density) of a closed triangle mesh.

    Signed-tetrahedron integration (Tonon, 'Explicit exact formulas for the 3-D tetrahedron inertia tensor', 2004).
    """
    a = verts[faces[:, 0]]
    b = verts[faces[:, 1]]
    c = verts[faces[:, 2]]
    vol6 = np.einsum("ij,ij->i", a, np.cross(b, c))
    volume = vol6.sum() / 6.0
    if abs(volume) < 1e-18:
        return 0.0, verts.mean(axis=0), np.zeros((3, 3))
    com = ((a + b + c) * vol6[:, None]).sum(axis=0) / (24.0 * volume)
    # Second moments about the origin: integral of x_i x_j over each signed tetrahedron (0, a, b, c).
    S = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            S[i, j] = (
                vol6
                * (
                    2 * a[:, i] * a[:, j]
                    + 2 * b[:, i] * b[:, j]
                    + 2 * c[:, i] * c[:, j]
                    + a[:, i] * b[:, j]
                    + a[:, j] * b[:, i]
                    + a[:, i] * c[:, j]
                    + a[:, j] * c[:, i]
                    + b[:, i] * c[:, j]
                    + b[:, j] * c[:, i]
                )
            ).sum() / 120.0
    S_com = S - volume * np.outer(com, com)
    inertia = np.trace(S_com) * np.eye(3) - S_com
    if volume < 0:
        volume, inertia = -volume, -inertia
    return float(volume), com, inertia


def principal_frame(inertia: np.ndarray) -> np.ndarray:
    """Right-handed rotation whose columns are the principal axes of ``inertia`` (eigenvalues ascending)."""
    _, vecs = np.linalg.eigh(0.5 * (inertia + inertia.T))
    if np.linalg.det(vecs) < 0:
        vecs[:, 2] = -vecs[:, 2]
    return vecs


def box_mesh(extents) -> tuple[np.ndarray, np.ndarray]:
    """Axis-aligned box centred at the origin (8 vertices, 12 outward triangles)."""
    hx, hy, hz = 0.5 * np.asarray(extents, dtype=np.float64)
    v = np.array(
        [[sx * hx, sy * hy, sz * hz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], dtype=np.float64
    )
    return convex_hull(v)


def capsule_points(radius: float, length: float, n_ring: int = 12, n_lat: int = 4) -> np.ndarray:
    """Point cloud of a z-aligned capsule (two hemispheres joined by a cylinder), used for AABBs and centres."""
    pts = []
    for sgn in (-1.0, 1.0):
        for k in range(n_lat + 1):
            el = 0.5 * np.pi * k / n_lat
            r = radius * np.cos(el)
            z = sgn * (0.5 * length + radius * np.sin(el))
            if r < 1e-12:
                pts.append([0.0, 0.0, z])
                continue
            for t in range(n_ring):
                az = 2 * np.pi * t / n_ring
                pts.append([r * np.cos(az), r * np.sin(az), z])
    return np.asarray(pts, dtype=np.float64)


def sphere_points(radius: float, n_ring: int = 12, n_lat: int = 6) -> np.ndarray:
    pts = [[0.0, 0.0, radius], [0.0, 0.0, -radius]]
    for k in range(1, n_lat):
        el = np.pi * k / n_lat - 0.5 * np.pi
        for t in range(n_ring):
            az = 2 * np.pi * t / n_ring
            pts.append([radius * np.cos(el) * np.cos(az), radius * np.cos(el) * np.sin(az), radius * np.sin(el)])
    return np.asarray(pts, dtype=np.float64)
