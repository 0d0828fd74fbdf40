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


# translation units: the C ABI with the per-stage kernels, and one unit per fused-kernel class (compiled in parallel)
SOURCES = ["rs_lib.hip", "rs_fused_t16.hip", "rs_fused_t24.hip", "rs_fused_gen.hip", "rs_fused_multi.hip"]
OBJ_DIR = os.path.join(HERE, "_build")


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile every translation unit (always, even when the library exists) and link librigidstep.so in place."""
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(OBJ_DIR, exist_ok=True)
    compile_flags = [f for f in FLAGS if f != "-shared"]

    def compile_one(src):
        obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
        cmd = ["hipcc", *compile_flags, "-c", "-o", obj, os.path.join(CSRC, src)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        return src, obj, r

    jobs = jobs or min(len(SOURCES), max(1, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        results = list(pool.map(compile_one, SOURCES))
    for src, _, r in results:
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed on {src} ({r.returncode})")
        if verbose:
            sys.stderr.write(r.stderr)
    tmp = OUT + ".tmp"
    cmd = ["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *[obj for _, obj, _ in results]]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc link failed ({r.returncode})")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
