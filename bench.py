"""Benchmark: env-steps/s of the batched rigid step on Franka x 4096 (one MI355X), fused hipGraph path vs plain.

Workload (one "step" = one scene.step() of all environments): plane + Franka Panda (9 DoF, convex-hull collision,
self-collision on), dt = 0.01 s, FP32, 4096 environments, initial joint positions uniform in the joint limits
(PCG64 seed 0), PD position targets resampled uniformly in the limits every 100 steps (PCG64 seed 1000 + k).
All inputs are resident in HBM before the timed region; targets are swapped in with device copies.

Prints one JSON line (rank 0). ``--gpus N`` with N > 1 is not supported: this build runs on one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np
import torch

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

from genesis_amd import runtime as rt  # noqa: E402
from genesis_amd.scenes import franka_initial_qpos, franka_model, franka_targets  # noqa: E402

HBM_PEAK_GBPS = 8000.0


def algorithmic_bytes_per_env_step(model, mean_overlap_pairs):
    """Persistent per-env state read+written once per step (FP32/int32), Franka in POSITION mode:
    qpos r+w 8Q, vel r+w 8D, acc w 4D, qacc_ws r+w 8D, ctrl_mode/ctrl_pos/ctrl_vel r 12D, errno r+w 8,
    n_contacts w 4, is_warmstart r+w 2, plus the warm-start normal cache r+w (24 B) of each AABB-overlapping pair."""
    Q, D = model.n_qs, model.n_dofs
    return 8 * Q + 8 * D + 4 * D + 8 * D + 12 * D + 8 + 4 + 2 + 24 * mean_overlap_pairs


def make_slab(model, dm, B):
    slab = rt.StateSlab(dm, B)
    slab["qpos"].copy_(torch.as_tensor(franka_initial_qpos(model, B).T, dtype=torch.float32))
    slab["ctrl_mode"].fill_(0)
    slab.forward_kinematics()
    return slab


def run_path(model, dm, B, steps, warmup, mode, targets, use_graph, block=100):
    """Returns (elapsed_s, per-step ms list, mean n_broad, slab)."""
    slab = make_slab(model, dm, B)
    stream = torch.cuda.Stream()
    n_blocks = len(targets)
    with torch.cuda.stream(stream):
        graph = rt.StepGraph(slab, block, mode, stream) if use_graph else None

        def do_block(k, n):
            slab["ctrl_pos"].copy_(targets[k % n_blocks])
            if graph is not None and n == block:
                graph.launch()
            else:
                slab.step(n, mode, stream)

        # warmup
        w = warmup
        k = 0
        while w > 0:
            n = min(block, w)
            do_block(k, n)
            w -= n
            k += 1
        stream.synchronize()
        # timed region
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        s = steps
        k = 0
        while s > 0:
            n = min(block, s)
            do_block(k, n)
            s -= n
            k += 1
        ev1.record(stream)
        stream.synchronize()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        gpu_ms = ev0.elapsed_time(ev1)
        # per-step latency: one substep at a time, each bracketed by events on the stream
        lat = []
        for i in range(min(200, steps)):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            slab.step(1, mode, stream)
            e1.record(stream)
            e1.synchronize()
            lat.append(e0.elapsed_time(e1))
        nb = slab["n_broad"].float().mean().item()
    return wall, gpu_ms, lat, nb, slab


def hbm_copy_peak_gbps():
    n = 1 << 28  # 1 GiB of float32
    a = torch.empty(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / 10
    del a, b
    return 2 * n * 4 / (ms * 1e-3) / 1e9


def cpu_baseline(model, n_envs=128, steps=200):
    """CPU restatement (FP32, same step) on a bounded sample of the workload: n_envs x steps, 1 thread."""
    from tests.oracle_lib import Oracle

    o = Oracle(model, n_envs, precision=32)
    o.set("qpos", franka_initial_qpos(model, n_envs))
    o.set("ctrl_mode", np.zeros((n_envs, model.n_dofs)))
    o.set("ctrl_pos", franka_targets(model, n_envs, 0))
    o.forward_kinematics()
    t0 = time.perf_counter()
    o.step(steps, 1)
    dt1 = time.perf_counter() - t0
    nthreads = min(os.cpu_count() or 1, 16)
    o2 = Oracle(model, n_envs, precision=32)
    o2.set("qpos", franka_initial_qpos(model, n_envs))
    o2.set("ctrl_mode", np.zeros((n_envs, model.n_dofs)))
    o2.set("ctrl_pos", franka_targets(model, n_envs, 0))
    o2.forward_kinematics()
    t0 = time.perf_counter()
    o2.step(steps, nthreads)
    dtn = time.perf_counter() - t0
    return n_envs * steps / dt1, n_envs * steps / dtn, nthreads


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--envs", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; this build measures one GPU "
              "(environments are independent; see DESIGN.md for how the path shards)", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))
    torch.cuda.set_device(0)
    B = args.envs
    model = franka_model()
    dm = rt.DeviceModel(model)
    n_blocks = max(1, (args.steps + args.warmup + 99) // 100)
    targets = [torch.as_tensor(franka_targets(model, B, k).T, dtype=torch.float32, device="cuda").contiguous()
               for k in range(n_blocks)]
    print(f"[bench] model D={model.n_dofs} G={model.n_geoms} pairs={model.n_pairs}; B={B}", file=sys.stderr,
          flush=True)
    wall_f, gpu_ms_f, lat_f, nb_f, slab = run_path(model, dm, B, args.steps, args.warmup, rt.RS_MODE_FUSED,
                                                   targets, use_graph=True)
    errno = slab.errno()
    print(f"[bench] fused: {wall_f:.3f}s for {args.steps} steps", file=sys.stderr, flush=True)
    plain = None
    if not args.no_plain:
        wall_p, gpu_ms_p, lat_p, nb_p, _ = run_path(model, dm, B, args.steps, args.warmup, rt.RS_MODE_PLAIN,
                                                    targets, use_graph=False)
        plain = {"value": B * args.steps / wall_p, "ms_per_step": 1e3 * wall_p / args.steps,
                 "p50_ms": float(np.percentile(lat_p, 50)), "p99_ms": float(np.percentile(lat_p, 99))}
        print(f"[bench] plain: {wall_p:.3f}s", file=sys.stderr, flush=True)
    value = B * args.steps / wall_f
    bytes_per_env_step = algorithmic_bytes_per_env_step(model, nb_f)
    kernel_ms = gpu_ms_f / args.steps  # one fused launch per step
    achieved = bytes_per_env_step * B / (kernel_ms * 1e-3) / 1e9
    out = {
        "metric": "env-steps/s at Franka×4096 (fused vs plain GPU, FP32); p50/p99 step latency",
        "value": value,
        "unit": "env-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall_f / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"Franka Panda x{B}, plane, dt=0.01, PD targets resampled every 100 steps",
                   "n_envs": B, "impl": "fused substep kernel, 100-step hipGraph"},
        "p50_step_ms": float(np.percentile(lat_f, 50)),
        "p99_step_ms": float(np.percentile(lat_f, 99)),
        "plain": plain,
        "fused_over_plain": (value / plain["value"]) if plain else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": None,
                     "bytes_per_env_step": bytes_per_env_step, "mean_overlap_pairs": nb_f,
                     "kernel": "k_fused_substep", "kernel_ms": kernel_ms},
        "errno": errno,
    }
    try:
        out["hbm_copy_GBps"] = hbm_copy_peak_gbps()
    except Exception as e:  # noqa: BLE001
        out["hbm_copy_GBps"] = None
        print(f"[bench] copy peak failed: {e}", file=sys.stderr)
    if not args.no_cpu_baseline and rank == 0:
        v1, vn, nt = cpu_baseline(model)
        out["cpu_baseline"] = {"value": v1, "unit": "env-steps/s", "cores": 1, "kind": "port",
                               "sample": "Franka x128 envs x 200 steps, same seeds, FP32 C++ restatement",
                               "value_all_cores": vn, "cores_all": nt, "cpu": platform.processor() or platform.machine()}
    if rank == 0:
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
