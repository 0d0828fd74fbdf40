"""Compile librigidstep.so for gfx950 with hipcc (in-tree, next to this file)."""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "librigidstep.so")
ARCH = os.environ.get("RS_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions are disabled on purpose (wrong results have been observed on this hardware)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # no FMA contraction: every product is rounded before it is added, as in the reference's sequential sums, so that
    # the per-stage and fused kernels (whose loops are shaped differently) compute bit-identical values
    "-ffp-contract=off",
]


def build(verbose: bool = False) -> str:
    src = os.path.join(CSRC, "rs_lib.hip")
    tmp = OUT + ".tmp"
    cmd = ["hipcc", *FLAGS, "-o", tmp, src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed ({r.returncode})")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
