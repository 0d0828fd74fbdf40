"""Helpers that move state between the device slab and the CPU restatement and compare fields."""

from __future__ import annotations

import numpy as np
import torch

# device slab field name -> oracle field name, where they differ
RENAME = {"errbits": "errno"}


def slab_to_oracle(slab, oracle, names=None):
    """Copy every field both sides know from the device slab into the oracle (env-major on the oracle side)."""
    for name, t in slab.fields.items():
        oname = RENAME.get(name, name)
        if names is not None and name not in names:
            continue
        try:
            rows = oracle.rows(oname)
        except KeyError:
            continue
        arr = t.detach().double().cpu().numpy()  # [rows, B]
        if arr.shape[0] != rows:
            continue
        oracle.set(oname, arr.T)


def oracle_to_slab(oracle, slab, names):
    for name in names:
        oname = RENAME.get(name, name)
        v = oracle.get(oname)  # [B, rows]
        t = slab.fields[name]
        t.copy_(torch.as_tensor(v.T, dtype=t.dtype, device=t.device))


def get_dev(slab, name):
    return slab.fields[name].detach().double().cpu().numpy().T  # [B, rows]


def rel_err(a, b, floor=1e-6):
    """max |a-b| / max(|b|, floor) per element, reduced over everything."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))) if a.size else 0.0


def scaled_err(a, b):
    """max |a-b| / (max |b| + tiny): error relative to the field's magnitude."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b)) / (np.max(np.abs(b)) + 1e-30))


# Models above this many dofs run the fused solver's Hessian, Cholesky factor and row products (J v, J^T f) on the
# matrix cores (rs_wave_solver.h wave_hessian_mfma, cholesky_blocked_mfma, gemv_rows_mfma, gemv_t_mfma), which sum in
# their own order: their fused step equals the per-stage step to the solver's tolerance, not to the last bit.
MFMA_MIN_DOFS = 25


SOLVER_TOL = 2e-3
_SOLVER_FIELDS = ("acc", "qacc_ws", "qf_constraint")
_VELOCITY_FIELDS = ("vel", "cd_vel")
_POSE_FIELDS = ("qpos", "links_pos", "links_quat", "geoms_pos", "aabb_min", "aabb_max")


def check_fused_matches_plain(plain, fused, model, names=("qpos", "vel", "acc", "n_contacts", "n_con")):
    """After one substep from identical state: bit-equal for models below MFMA_MIN_DOFS. Above it, the contact and
    row counts equal (detection precedes the solve), and the solver's outputs (constrained acceleration, its warm
    start, the constraint force in joint space) agree within SOLVER_TOL of s = max(|acc|, 1) per environment: the
    solver stops at its tolerance (solver.py:3063-3139), so two Hessians or row products that differ in the last bits
    give iterates that differ up to it. What the integrator derives from acc carries that difference on:
    vel' = vel + dt acc moves by dt SOLVER_TOL s, and the poses (qpos' = qpos + dt vel', the bodies' frames, which
    follow qpos for the free bodies of these models) by dt^2 SOLVER_TOL s, each on top of 5e-5 of the field's
    magnitude; every other field within 5e-5 of its magnitude."""
    dt = model.options.dt
    s_env = None
    for name in names:
        a, b = get_dev(plain, name), get_dev(fused, name)
        if model.n_dofs < MFMA_MIN_DOFS or name in ("n_contacts", "n_con", "n_broad", "errbits"):
            d = float(np.max(np.abs(a - b))) if a.size else 0.0
            assert d == 0.0, f"{name}: fused differs from per-stage by {d:.3e}"
            continue
        if s_env is None:
            s_env = np.maximum(np.abs(get_dev(plain, "acc")).max(axis=1, keepdims=True), 1.0)
        a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
        if name in _SOLVER_FIELDS or name in _VELOCITY_FIELDS or name in _POSE_FIELDS:
            k = 1.0 if name in _SOLVER_FIELDS else dt if name in _VELOCITY_FIELDS else dt * dt
            bound = (0.0 if name in _SOLVER_FIELDS else 5e-5 * np.max(np.abs(a))) + k * SOLVER_TOL * s_env
            ratio = float(np.max(np.abs(b - a) / bound))
            assert ratio < 1.0, f"{name}: {ratio:.3f} of the bound"
        else:
            err = scaled_err(b, a)
            assert err < 5e-5, f"{name}: {err:.3e}"


def check_fused_as_accurate_as_plain(plain, fused, oracle64, model, names=("qpos", "vel", "acc")):
    """For a solver that stops at its iteration cap before its tolerance (the conjugate-gradient solver on 120 dofs,
    solver.py:3063-3139 with iterations=50), last-bit differences in the row products move the unconverged iterate by
    more than the tolerance, so the bar is accuracy instead: after one substep from identical state, the fused step's
    distance from the FP64 restatement (per environment, relative to max(|x|, 1)) is at most 1.5 times the per-stage
    step's own plus 1e-6, and the contact and row counts equal."""
    for name in ("n_contacts", "n_con"):
        assert np.array_equal(get_dev(plain, name), get_dev(fused, name)), name
    for name in names:
        a, b, o = (np.asarray(x, np.float64) for x in (get_dev(plain, name), get_dev(fused, name), oracle64.get(name)))
        scale = np.maximum(np.abs(o).max(axis=1, keepdims=True), 1.0)
        ea = float(np.max(np.abs(a - o) / scale))
        eb = float(np.max(np.abs(b - o) / scale))
        print(f"[fused] {name}: per-stage {ea:.3e}, fused {eb:.3e} from the FP64 restatement", flush=True)
        assert eb <= 1.5 * ea + 1e-6, f"{name}: fused {eb:.3e} vs per-stage {ea:.3e}"
