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


# Models above this many dofs assemble the fused solver's Hessian on the matrix cores (rs_wave_solver.h
# wave_hessian_mfma), which sum the rows in their own order: their fused step equals the per-stage step to the solver's
# tolerance, not to the last bit.
MFMA_MIN_DOFS = 25


def check_fused_matches_plain(plain, fused, model, names=("qpos", "vel", "acc", "n_contacts", "n_con")):
    """After one substep from identical state: bit-equal for models below MFMA_MIN_DOFS; above it, the contact and
    row counts equal (detection precedes the solve), qpos/vel within 5e-5 of their magnitude, and the constrained
    acceleration within 2e-3 of max(|acc|, 1) per environment (the solver stops at its tolerance, solver.py:3063-3139:
    two Hessians that differ in the last bits give iterates that differ up to it)."""
    for name in names:
        a, b = get_dev(plain, name), get_dev(fused, name)
        if model.n_dofs < MFMA_MIN_DOFS or name in ("n_contacts", "n_con", "n_broad"):
            d = float(np.max(np.abs(a - b))) if a.size else 0.0
            assert d == 0.0, f"{name}: fused differs from per-stage by {d:.3e}"
        elif name == "acc":
            scale = np.maximum(np.abs(a).max(axis=1, keepdims=True), 1.0)
            err = float(np.max(np.abs(b - a) / scale))
            assert err < 2e-3, f"acc: {err:.3e}"
        else:
            err = scaled_err(b, a)
            assert err < 5e-5, f"{name}: {err:.3e}"
