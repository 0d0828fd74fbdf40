"""rs_errno_reduce_noalloc (the allocation-free variant of rs_errno_reduce) against rs_errno_reduce, through the C ABI.

Two calls in a row on different slabs, the second smaller, each compared with the allocating entry point; then a clean
slab after a dirty one (the static scratch keeps the last call's contents, so a stale word would show here), and the
rejection of an environment count beyond the static scratch.
"""

from __future__ import annotations

import ctypes

import pytest
import torch

from genesis_amd import scenes

pytestmark = pytest.mark.gpu


def _slab(rt, dm, model, B, nan_envs):
    slab = rt.StateSlab(dm, B)
    slab["qpos"].copy_(torch.as_tensor(scenes.franka_initial_qpos(model, B).T, dtype=torch.float32))
    slab["ctrl_mode"].fill_(2)
    for b in nan_envs:
        slab["vel"][1, b] = float("nan")
    slab.forward_kinematics()
    slab.step(1, rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    return slab


def _both(rt, slab, n_envs=None):
    lib = rt.load_library()
    n = slab.n_envs if n_envs is None else n_envs
    a, b = ctypes.c_int32(-1), ctypes.c_int32(-1)
    rt._check(lib.rs_errno_reduce(slab.dmodel.handle, slab.ptr, n, ctypes.byref(a), None))
    rc = lib.rs_errno_reduce_noalloc(slab.dmodel.handle, slab.ptr, n, ctypes.byref(b), None)
    return a.value, b.value, rc


def test_noalloc_errno_equals_the_allocating_entry_point():
    from genesis_amd import runtime as rt

    rt.load_library()
    model = scenes.franka_model()
    dm = rt.DeviceModel(model)
    # 600 environments: three blocks of partial results; NaNs in the first and the last block
    big = _slab(rt, dm, model, 600, nan_envs=(3, 590))
    print("[noalloc] call 1: 600 envs, NaN in envs 3 and 590", flush=True)
    a, b, rc = _both(rt, big)
    assert rc == rt.RS_OK
    assert a == b == (rt.ERR_INVALID_FORCE_NAN | rt.ERR_INVALID_ACC_NAN), (a, b)
    # a smaller slab next (one block), its NaN in the last environment
    small = _slab(rt, dm, model, 70, nan_envs=(69,))
    print("[noalloc] call 2: 70 envs, NaN in env 69", flush=True)
    a, b, rc = _both(rt, small)
    assert rc == rt.RS_OK and a == b == (rt.ERR_INVALID_FORCE_NAN | rt.ERR_INVALID_ACC_NAN), (a, b)
    # a clean slab after the dirty ones: no word of the earlier calls may leak into the result
    clean = _slab(rt, dm, model, 300, nan_envs=())
    print("[noalloc] call 3: 300 clean envs", flush=True)
    a, b, rc = _both(rt, clean)
    assert rc == rt.RS_OK and a == b == 0, (a, b)
    # the first slab again, over its first 64 environments only (the NaN in env 3 is inside, 590 outside)
    a, b, rc = _both(rt, big, n_envs=64)
    assert rc == rt.RS_OK and a == b == (rt.ERR_INVALID_FORCE_NAN | rt.ERR_INVALID_ACC_NAN), (a, b)


def test_noalloc_errno_rejects_counts_beyond_its_scratch():
    from genesis_amd import runtime as rt

    lib = rt.load_library()
    model = scenes.franka_model()
    dm = rt.DeviceModel(model)
    slab = rt.StateSlab(dm, 4)
    out = ctypes.c_int32(-7)
    # checked before anything is launched: the state pointer is never read for this count
    rc = lib.rs_errno_reduce_noalloc(dm.handle, slab.ptr, (1 << 20) + 1, ctypes.byref(out), None)
    assert rc == rt.RS_E_INVALID
    assert b"at most 1048576 environments" in lib.rs_last_error()
    assert out.value == -7
