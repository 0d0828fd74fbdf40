"""MI355X-native batched rigid-body step with the reference's ``gs.*`` scene API.

``import genesis_amd as gs`` then ``gs.init(backend=gs.gpu)``, ``gs.Scene(...)``, ``scene.add_entity(gs.morphs.MJCF(...))``,
``scene.build(n_envs=N)``, ``scene.step()`` (genesis/__init__.py:57-300, genesis/engine/scene.py).
"""

from __future__ import annotations

from . import _global_state, materials, morphs, options
from .constants import (
    CTRL_MODE,
    EQUALITY_TYPE,
    GEOM_TYPE,
    JOINT_TYPE,
    backend,
    broadphase_traversal,
    constraint_solver,
    integrator,
)
from .scene import GenesisException, Scene

gpu = backend.gpu
amdgpu = backend.amdgpu
cpu = backend.cpu
cuda = backend.cuda

__version__ = "0.1.0"


def init(backend=None, precision: str = "32", seed: int | None = None, performance_mode: bool = False,
         debug: bool = False, logging_level=None, **_ignored):
    """Select the device (genesis/__init__.py:57). Only FP32 on a ROCm GPU is built."""
    _global_state.init(backend, precision, seed)


def destroy():
    _global_state.destroy()


def __getattr__(name):
    if name == "device":
        return _global_state.device()
    if name == "EPS":
        return 1.1920928955078125e-07
    raise AttributeError(name)
