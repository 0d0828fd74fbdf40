"""Morph descriptors (``gs.morphs``) for the shapes the rigid step supports.

Mirrors ``genesis/options/morphs.py``: ``Plane`` (``:427-468``), ``Box``, ``Sphere``, ``MJCF`` and ``URDF``
(``:916-1030``; file paths relative to an asset directory or absolute). Each morph turns into an entity spec for the
model compiler. Asset directories: the package's own ``assets`` and, when set, the directories listed in the
``GENESIS_AMD_ASSET_DIR`` environment variable (``os.pathsep``-separated), e.g. a Genesis checkout's
``genesis/assets``.
"""

from __future__ import annotations

import os

import numpy as np

from .model import primitives
from .model.mjcf import load_mjcf
from .model.urdf import DEFAULT_ARMATURE, load_urdf
from .model import transforms as tf

ASSET_DIRS = [os.path.join(os.path.dirname(os.path.abspath(__file__)), "assets")]


def asset_dirs() -> list[str]:
    extra = [d for d in os.environ.get("GENESIS_AMD_ASSET_DIR", "").split(os.pathsep) if d]
    return ASSET_DIRS + extra


def find_asset(path: str) -> str:
    if os.path.isabs(path) and os.path.exists(path):
        return path
    for d in asset_dirs() + [os.getcwd()]:
        p = os.path.join(d, path)
        if os.path.exists(p):
            return p
    raise FileNotFoundError(f"asset '{path}' not found in {asset_dirs()}")


class Morph:
    def __init__(self, pos=(0.0, 0.0, 0.0), quat=(1.0, 0.0, 0.0, 0.0), euler=None, fixed=False, visualization=True,
                 collision=True, **kwargs):
        self.pos = np.asarray(pos, dtype=np.float64)
        if euler is not None:
            self.quat = tf.euler_xyz_to_quat(np.deg2rad(np.asarray(euler, dtype=np.float64)))
        else:
            self.quat = tf.quat_normalize(quat)
        self.fixed = fixed
        self.visualization = visualization
        self.collision = collision
        self.extra = kwargs


class Plane(Morph):
    def __init__(self, normal=(0.0, 0.0, 1.0), plane_size=(1e3, 1e3), **kwargs):
        super().__init__(fixed=True, **kwargs)
        self.normal = normal
        self.plane_size = plane_size

    def to_spec(self, material, name):
        return primitives.plane_entity(self.normal, self.plane_size, friction=_friction(material, 1.0), name=name)


class Box(Morph):
    def __init__(self, size=(1.0, 1.0, 1.0), lower=None, upper=None, **kwargs):
        super().__init__(**kwargs)
        if lower is not None and upper is not None:
            lower, upper = np.asarray(lower, float), np.asarray(upper, float)
            self.size = upper - lower
            self.pos = 0.5 * (upper + lower)
        else:
            self.size = np.asarray(size, dtype=np.float64)

    def to_spec(self, material, name):
        return primitives.box_entity(self.size, self.pos, self.quat, rho=_rho(material),
                                     friction=_friction(material, 1.0), fixed=self.fixed, name=name)


class Sphere(Morph):
    def __init__(self, radius=0.5, **kwargs):
        super().__init__(**kwargs)
        self.radius = radius

    def to_spec(self, material, name):
        return primitives.sphere_entity(self.radius, self.pos, self.quat, rho=_rho(material),
                                        friction=_friction(material, 1.0), fixed=self.fixed, name=name)


class MJCF(Morph):
    def __init__(self, file: str, scale: float = 1.0, **kwargs):
        super().__init__(**kwargs)
        if scale != 1.0:
            raise NotImplementedError("MJCF scale != 1 is not supported")
        self.file = file

    def to_spec(self, material, name):
        spec = load_mjcf_or_cached(self.file, name)
        root = spec.links[0]
        if not np.allclose(self.pos, 0) or not np.allclose(self.quat, [1, 0, 0, 0]):
            for link in spec.links:
                if link.parent < 0:
                    link.pos, link.quat = tf.compose(self.pos, self.quat, link.pos, link.quat)
                    for j in link.joints:
                        if j.init_qpos is not None and len(j.init_qpos) == 7:
                            j.init_qpos = np.concatenate([link.pos, link.quat])
        if material is not None and material.friction is not None:
            for link in spec.links:
                for g in link.geoms:
                    g.friction = material.friction
        del root
        return spec


class URDF(Morph):
    """``options/morphs.py:916-1030``: ``fixed`` (default False: the base link floats on a free joint),
    ``merge_fixed_links`` / ``links_to_keep``, ``default_armature``; ``pos``/``quat``/``euler`` place the base."""

    def __init__(self, file: str, scale: float = 1.0, merge_fixed_links: bool = True, links_to_keep=(),
                 default_armature: float | None = DEFAULT_ARMATURE, fixed: bool = False, **kwargs):
        super().__init__(fixed=fixed, **kwargs)
        if scale != 1.0:
            raise NotImplementedError("URDF scale != 1 is not supported")
        self.file = file
        self.merge_fixed_links = merge_fixed_links
        self.links_to_keep = tuple(links_to_keep)
        self.default_armature = default_armature

    def to_spec(self, material, name):
        spec = load_urdf(find_asset(self.file), name, fixed=self.fixed, merge_fixed=self.merge_fixed_links,
                         links_to_keep=self.links_to_keep, default_armature=self.default_armature, pos=self.pos,
                         quat=self.quat)
        if material is not None and material.friction is not None:
            for link in spec.links:
                for g in link.geoms:
                    g.friction = material.friction
        return spec


SPEC_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "assets", "specs")


def cached_spec_path(file: str) -> str:
    """Saved loader output for an MJCF asset path (used where the MJCF source tree is absent)."""
    key = os.path.normpath(file).replace(os.sep, "__")
    return os.path.join(SPEC_DIR, os.path.splitext(key)[0] + ".npz")


def load_mjcf_or_cached(file: str, name: str):
    from .model.spec_io import load_spec

    try:
        path = find_asset(file)
    except FileNotFoundError:
        cached = cached_spec_path(file)
        if not os.path.exists(cached):
            raise
        spec = load_spec(cached)
        spec.name = name
        return spec
    return load_mjcf(path, name)


def _rho(material):
    return 1000.0 if material is None or material.rho is None else material.rho


def _friction(material, default):
    return default if material is None or material.friction is None else material.friction
