"""Process-wide init state (backend, precision, seed), like ``gs.init`` (genesis/__init__.py:57-300)."""

from __future__ import annotations

import torch

from .constants import backend as _backend

_state = {"initialized": False, "backend": None, "precision": "32", "seed": None}


def init(backend=None, precision="32", seed=None):
    if _state["initialized"]:
        raise RuntimeError("genesis_amd is already initialized")
    if precision != "32":
        raise NotImplementedError("only precision='32' is built on the device (the FP64 oracle is test-only)")
    if backend not in (None, _backend.gpu, _backend.amdgpu):
        raise NotImplementedError(f"backend {backend!r} is not built: the rigid step runs on a ROCm GPU only")
    if not torch.cuda.is_available():
        raise RuntimeError("genesis_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    from . import runtime

    runtime.load_library()
    if seed is not None:
        torch.manual_seed(seed)
    _state.update(initialized=True, backend=_backend.gpu, precision=precision, seed=seed)


def destroy():
    _state.update(initialized=False, backend=None)


def require_init():
    if not _state["initialized"]:
        raise RuntimeError("genesis_amd is not initialized: call gs.init() first")


def device():
    return torch.device("cuda", torch.cuda.current_device())
