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
