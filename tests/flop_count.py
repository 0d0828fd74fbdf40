"""Floating-point operations per env-step of each bench workload, counted by the restatement instantiated on an
operation-counting double (oracle/oracle_flops.h: +, -, *, / and sqrt one each, library calls one each and also
counted apart, comparisons and fabs free). The count follows the bench's own run: the same seeded initial states,
the same PD targets resampled every 100 substeps, the warm-up then the timed substeps, on a sample of the
environments; every 10th substep runs stage by stage to split the count by stage. The device kernels run the same
arithmetic in the same order (no FMA contraction), so the count is the device step's FP work.

Writes profiles/flops_<workload>.json, which bench.py reads for its VALU roofline line. Test infrastructure (it runs
the restatement); run it as ``python -m tests.flop_count [--workload franka|boxpile|quadruped|all]``.
"""

from __future__ import annotations

import argparse
import json
import os

import numpy as np

from genesis_amd.runtime import STAGE_ID
from tests.oracle_lib import Oracle, flops

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STEP_STAGES = ("crba", "factor", "rne", "aabb", "broadphase", "narrowphase", "constraints", "solve", "integrate")
# (sampled environments, warm-up substeps, timed substeps) of each workload: bench.py's defaults, fewer environments
SAMPLES = {"franka": (64, 100, 1000), "boxpile": (16, 100, 500), "quadruped": (64, 100, 1000)}
OUT_NAMES = {"franka": "flops_franka4096.json", "boxpile": "flops_boxpile1024.json",
             "quadruped": "flops_quadruped8192.json"}


def count_workload(name: str) -> dict:
    import bench

    wl = bench.Workload(name)
    n_envs, warmup, steps = SAMPLES[name]
    o = Oracle(wl.model, n_envs, precision=0)
    bench._oracle_init(wl, o, n_envs)
    total = warmup + steps
    per_stage = {s: 0 for s in STEP_STAGES}
    staged_steps = 0
    timed_ops = timed_tr = 0
    for t in range(total):
        if t % bench.BLOCK == 0 and t > 0:
            tg = wl.targets(n_envs, t // bench.BLOCK)
            if tg is not None:
                o.set("ctrl_pos", tg.astype(np.float32).astype(np.float64))
        f0 = flops()
        if t >= warmup and t % 10 == 0:
            for s in STEP_STAGES:
                a = flops()[0]
                o.stage(STAGE_ID[s])
                per_stage[s] += flops()[0] - a
            staged_steps += 1
        else:
            o.step(1, 1)
        f1 = flops()
        if t >= warmup:
            timed_ops += f1[0] - f0[0]
            timed_tr += f1[1] - f0[1]
    env_steps = n_envs * steps
    return {
        "workload": wl.desc.format(B=wl.B),
        "sample": f"{n_envs} of the bench's environments, {warmup} warm-up + {steps} timed substeps; stage split "
                  f"from every 10th timed substep ({staged_steps} substeps)",
        "counting": "binary +, -, *, / and sqrt: 1 each; sin, cos, atan2, acos, pow, fmod: 1 each (also counted apart); "
                    "negation, fabs, comparisons, fmin/fmax, floor/ceil: 0",
        "flops_per_env_step": timed_ops / env_steps,
        "library_calls_per_env_step": timed_tr / env_steps,
        "per_stage_flops_per_env_step": {s: v / (n_envs * max(staged_steps, 1)) for s, v in per_stage.items()},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="all", choices=["franka", "boxpile", "quadruped", "all"])
    args = ap.parse_args()
    names = list(SAMPLES) if args.workload == "all" else [args.workload]
    for name in names:
        res = count_workload(name)
        path = os.path.join(REPO, "profiles", OUT_NAMES[name])
        with open(path, "w") as f:
            json.dump(res, f, indent=1)
        print(f"{name}: {res['flops_per_env_step']:.4g} FLOP per env-step -> {path}", flush=True)


if __name__ == "__main__":
    main()
