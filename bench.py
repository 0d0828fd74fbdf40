"""Benchmark: env-steps/s of the batched rigid step on one MI355X, fused hipGraph path vs plain per-stage kernels.

Default workload (the headline metric): Franka Panda x 4096 - plane + bundled Panda MJCF (9 DoF, convex-hull
collision geoms, self-collision on), dt = 0.01 s, FP32, initial joint positions uniform in the joint limits (PCG64
seed 0), PD position targets resampled uniformly in the limits every 100 steps (PCG64 seed 1000 + k).
Other workloads (``--workload``): ``boxpile`` (20 free boxes x 1024 envs, dt 0.01) and ``quadruped`` (18-DoF
quadruped x 8192 envs, dt 0.005, PD targets around the standing pose).
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
from genesis_amd import scenes  # noqa: E402

HBM_PEAK_GBPS = 8000.0
FP32_VECTOR_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak (counts an FMA as 2 and packed FP32)
FLOP_FILES = {"franka": "flops_franka4096.json", "boxpile": "flops_boxpile1024.json",
              "quadruped": "flops_quadruped8192.json"}
BLOCK = 100  # substeps per captured graph (and per target block)


class Workload:
    def __init__(self, name):
        self.name = name
        if name == "franka":
            self.model = scenes.franka_model()
            self.B, self.steps = 4096, 1000
            self.desc = "Franka Panda x{B}, plane, dt=0.01, PD targets resampled every 100 steps"
        elif name == "boxpile":
            # box-box pairs through the separating-axis detection (box_box_detection=True), as the workload asks
            self.model = scenes.box_pile_model(box_box_detection=True)
            self.B, self.steps = 1024, 500
            self.desc = "box pile (20 free boxes, box-box SAT detection) x{B}, plane, dt=0.01"
        elif name == "quadruped":
            self.model = scenes.quadruped_model()
            self.B, self.steps = 8192, 1000
            self.desc = "quadruped (18 DoF) x{B}, plane, dt=0.005, PD targets around standing pose"
        else:
            raise SystemExit(f"unknown workload {name}")

    def init_state(self, slab, B):
        m = self.model
        if self.name == "franka":
            slab["qpos"].copy_(torch.as_tensor(scenes.franka_initial_qpos(m, B).T, dtype=torch.float32))
            slab["ctrl_mode"].fill_(0)
        elif self.name == "quadruped":
            slab["qpos"].copy_(torch.as_tensor(scenes.quadruped_initial_qpos(m, B).T, dtype=torch.float32))
            slab["ctrl_mode"].copy_(torch.as_tensor(scenes.quadruped_ctrl_mode(m, B).T, dtype=torch.int32))
        slab.forward_kinematics()

    def targets(self, B, k):
        m = self.model
        if self.name == "franka":
            return scenes.franka_targets(m, B, k)
        if self.name == "quadruped":
            return scenes.quadruped_targets(m, B, k)
        return None


def algorithmic_bytes_per_env_step(model, mean_overlap_pairs):
    """Persistent per-env state read+written once per step (FP32/int32): qpos r+w 8Q, vel r+w 8D, acc w 4D,
    qacc_ws r+w 8D, ctrl_mode/ctrl_pos/ctrl_vel r 12D, errno r+w 8, n_contacts w 4, is_warmstart r+w 2, plus the
    warm-start normal cache r+w (24 B) of each AABB-overlapping pair."""
    Q, D = model.n_qs, model.n_dofs
    return 8 * Q + 8 * D + 4 * D + 8 * D + 12 * D + 8 + 4 + 2 + 24 * mean_overlap_pairs


def valu_roofline(wl, kernel_ms, B):
    """FP32 work of the dominant kernel against the vector peak: the FLOPs per env-step of this workload's step,
    counted once by the restatement on an operation-counting double over the bench's own run (profiles/flops_*.json,
    written by tests/flop_count.py), times the environments one launch steps, over the launch's measured time."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", FLOP_FILES[wl.name])
    try:
        with open(path) as f:
            fl = json.load(f)
    except (OSError, ValueError):
        return None
    achieved = fl["flops_per_env_step"] * B / (kernel_ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": FP32_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP32_VECTOR_PEAK_TFLOPS, "flops_per_env_step": fl["flops_per_env_step"],
            "per_stage_flops_per_env_step": fl["per_stage_flops_per_env_step"],
            "source": f"profiles/{FLOP_FILES[wl.name]} ({fl['sample']})"}


def make_slab(wl, dm, B):
    slab = rt.StateSlab(dm, B)
    wl.init_state(slab, B)
    return slab


def run_path(wl, dm, B, steps, warmup, mode, targets, use_graph, n_lat=200):
    """Returns (wall_s, gpu_ms, per-substep latencies ms, mean n_broad, slab, timed-region launch counts, graph
    substeps).

    Targets change every BLOCK substeps (counted from the first warm-up substep). The fused path replays a hipGraph
    of min(BLOCK, steps) substeps, so the timed region runs the graph for any --steps that divides into it."""
    slab = make_slab(wl, dm, B)
    stream = torch.cuda.Stream()
    gblock = min(BLOCK, steps)
    with torch.cuda.stream(stream):
        graph = rt.StepGraph(slab, gblock, mode, stream) if use_graph else None
        state = {"t": 0, "graph_launches": 0, "eager_substeps": 0}

        def run(n_total):
            while n_total > 0:
                t = state["t"]
                n = min(gblock, n_total, BLOCK - t % BLOCK)
                if targets is not None and t % BLOCK == 0:
                    slab["ctrl_pos"].copy_(targets[(t // BLOCK) % len(targets)])
                if graph is not None and n == gblock:
                    graph.launch()
                    state["graph_launches"] += 1
                else:
                    slab.step(n, mode, stream)
                    state["eager_substeps"] += n
                state["t"] += n
                n_total -= n

        run(warmup)
        stream.synchronize()
        torch.cuda.synchronize()
        state["graph_launches"] = state["eager_substeps"] = 0
        t0 = time.perf_counter()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        run(steps)
        ev1.record(stream)
        stream.synchronize()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        gpu_ms = ev0.elapsed_time(ev1)
        timed = dict(state)
        # per-substep latency: one substep at a time, each bracketed by events on the launch stream
        lat = []
        for _ in range(min(n_lat, steps)):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            slab.step(1, mode, stream)
            e1.record(stream)
            e1.synchronize()
            lat.append(e0.elapsed_time(e1))
        nb = slab["n_broad"].float().mean().item()
    return wall, gpu_ms, lat, nb, slab, timed, gblock


COPY_BYTES = 1 << 30


def hbm_copy_peak_gbps():
    """Streaming copy of COPY_BYTES through torch's elementwise kernel (b = a * 1): the achievable-bandwidth figure,
    and the calibration dispatches that the FETCH_SIZE / WRITE_SIZE counter runs scale by."""
    n = COPY_BYTES // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        torch.mul(a, 1.0, out=b)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / 10
    del a, b
    return 2 * n * 4 / (ms * 1e-3) / 1e9


def measured_traffic(kernel, wl, B):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3 counter runs of this workload, if they were
    taken on the current kernel sources (same source digest); else None."""
    from genesis_amd.profiling import source_digest

    path = os.path.join(REPO, "profiles", f"{kernel}_{wl.name}{B}.json")
    if not os.path.exists(path):
        return None, "no counter profile for this workload"
    with open(path) as f:
        fig = json.load(f)
    if fig.get("source_digest") != source_digest():
        return None, f"{os.path.basename(path)} was taken on other kernel sources"
    return fig["traffic"]["bytes_per_launch"], os.path.basename(path)


def cpu_model_name():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def _oracle_init(wl, o, n_envs, block=0):
    m = wl.model
    if wl.name == "franka":
        o.set("qpos", scenes.franka_initial_qpos(m, n_envs).astype(np.float32).astype(np.float64))
        o.set("ctrl_mode", np.zeros((n_envs, m.n_dofs)))
    elif wl.name == "quadruped":
        o.set("qpos", scenes.quadruped_initial_qpos(m, n_envs).astype(np.float32).astype(np.float64))
        o.set("ctrl_mode", scenes.quadruped_ctrl_mode(m, n_envs))
    t = wl.targets(n_envs, block)
    if t is not None:
        o.set("ctrl_pos", t.astype(np.float32).astype(np.float64))
    o.forward_kinematics()


def cpu_baseline(wl):
    """The FP32 CPU restatement of the same step, timed on this host at 1 thread and at min(nproc, 16) threads, on
    a bounded sample: Franka x64 envs x 1000 steps with the targets resampled every 100 steps (the reference's
    CPU-sized configuration); box pile x32 x 100 steps; quadruped x64 x 500 steps."""
    from tests.oracle_lib import Oracle

    n_envs, steps = {"franka": (64, 1000), "boxpile": (32, 100), "quadruped": (64, 500)}[wl.name]
    out = []
    nthreads = min(os.cpu_count() or 1, 16)
    for threads in (1, nthreads):
        o = Oracle(wl.model, n_envs, precision=32)
        _oracle_init(wl, o, n_envs)
        dt = 0.0
        for k in range((steps + BLOCK - 1) // BLOCK):
            t = wl.targets(n_envs, k)
            if t is not None:
                o.set("ctrl_pos", t.astype(np.float32).astype(np.float64))
            n = min(BLOCK, steps - k * BLOCK)
            t0 = time.perf_counter()
            o.step(n, threads)
            dt += time.perf_counter() - t0
        out.append(n_envs * steps / dt)
    return out[0], out[1], nthreads, f"{wl.name} x{n_envs} envs x {steps} steps"


def drift_vs_fp64(wl, dm):
    """FP32 device trajectory against the FP64 CPU restatement (and the FP32 restatement against FP64, for scale) on
    a sample of the workload from the same seeded state, targets held at their first draw: max over dofs of
    |q32 - q64| per environment, summarised over environments."""
    from tests.oracle_lib import Oracle

    n_envs, steps = {"franka": (64, 1000), "boxpile": (32, 500), "quadruped": (64, 1000)}[wl.name]
    slab = make_slab(wl, dm, n_envs)
    t = wl.targets(n_envs, 0)
    if t is not None:
        slab["ctrl_pos"].copy_(torch.as_tensor(t.T, dtype=torch.float32))
    slab.step(steps, rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    q_dev = slab["qpos"].double().cpu().numpy().T
    res = {}
    for p in (64, 32):
        o = Oracle(wl.model, n_envs, precision=p)
        _oracle_init(wl, o, n_envs)
        o.step(steps, min(os.cpu_count() or 1, 16))
        res[p] = o.get("qpos")
    d = np.abs(q_dev - res[64]).max(axis=1)
    d32 = np.abs(res[32] - res[64]).max(axis=1)
    return {"sample": f"{wl.name} x{n_envs} envs x {steps} steps", "unit": "rad or m (max over qpos entries)",
            "device_fp32": {"median": float(np.median(d)), "p90": float(np.percentile(d, 90)), "max": float(d.max())},
            "cpu_fp32": {"median": float(np.median(d32)), "p90": float(np.percentile(d32, 90)),
                         "max": float(d32.max())}}


def like_for_like(n_envs=30000, steps=1000):
    """The reference's own speed benchmark (examples/speed_benchmark/franka.py: plane + Panda MJCF, dt 0.01,
    n_envs=30000, PD to the collision-free pose [0,0,0,-1,0,1,0,0.02,0.02] from the neutral pose, 1000 scene.step()
    calls), run through this library's Scene API (autotuned, graph-replayed fused step). The reference publishes
    43 M env-steps/s for it on one RTX 4090 (README.md:60, franka.py:38)."""
    import genesis_amd as gs

    gs.init(backend=gs.gpu)
    try:
        scene = gs.Scene(rigid_options=gs.options.RigidOptions(dt=0.01), show_viewer=False)
        scene.add_entity(gs.morphs.Plane())
        franka = scene.add_entity(gs.morphs.MJCF(file="xml/franka_emika_panda/panda.xml"))
        scene.build(n_envs=n_envs)
        franka.control_dofs_position(torch.tile(torch.tensor([0, 0, 0, -1.0, 0, 1.0, 0, 0.02, 0.02],
                                                             device="cuda"), (n_envs, 1)))
        for _ in range(20):  # autotune on the first step, graph capture, warm-up
            scene.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            scene.step()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return {"workload": f"examples/speed_benchmark/franka.py: Franka x{n_envs}, fixed PD target, "
                            f"{steps} scene.step() calls", "value": n_envs * steps / wall, "unit": "env-steps/s",
                "ms_per_step": 1e3 * wall / steps, "published": 43e6,
                "published_hardware": "1x NVIDIA RTX 4090 (README.md:60)",
                "ratio_to_published": n_envs * steps / wall / 43e6, "fused_config": scene.dmodel.fused_config}
    finally:
        gs.destroy()


def serving_bench(args):
    """Continuous-batching rollout server on a seeded Poisson mix: scene uniform over {Franka, box pile, quadruped},
    n_envs log-uniform in [64, 4096], horizon uniform in [50, 500] (seed 4). Reports aggregate env-steps/s, the
    per-tick (one shared fused launch) p50/p99 latency, and whether three requests replayed alone end bit-identical
    to their run in the mix."""
    from genesis_amd.serving import RolloutServer, poisson_requests

    reqs = poisson_requests(args.requests, rate_hz=args.rate, seed=4, min_envs=64, max_envs=4096, min_horizon=50,
                            max_horizon=500)
    server = RolloutServer(hbm_budget_bytes=64 << 30, max_active=32)
    warm = poisson_requests(3, rate_hz=1e6, seed=5, min_envs=64, max_envs=256, min_horizon=5, max_horizon=10)
    server.run(warm)  # model uploads and first launches outside the measurement
    server.tick_ms.clear()
    server.active_per_tick.clear()
    print(f"[bench] serving: {len(reqs)} requests, rate {args.rate}/s", file=sys.stderr, flush=True)
    stats = server.run(reqs, keep_final=True)
    equal = []
    for r in sorted(reqs, key=lambda r: r.n_envs)[:3]:
        slab = server.new_slab(r)
        slab.step(r.horizon, rt.RS_MODE_FUSED)
        torch.cuda.synchronize()
        equal.append(bool(torch.equal(slab["qpos"], r.final_qpos) and torch.equal(slab["vel"], r.final_vel)))
    out = {
        "metric": "serving mix: aggregate env-steps/s; p50/p99 shared-launch step latency",
        "value": stats["env_steps_per_s"],
        "unit": "env-steps/s",
        "n_gpus": 1,
        "steps": stats["ticks"],
        "warmup": 0,
        "ms_per_step": 1e3 * stats["wall_s"] / max(1, stats["ticks"]),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"Poisson mix of {len(reqs)} requests at {args.rate}/s (seed 4): scene uniform over "
                               "franka/boxpile/quadruped, n_envs log-U[64,4096], horizon U[50,500]"},
        "p50_step_ms": stats["p50_step_ms"],
        "p99_step_ms": stats["p99_step_ms"],
        "serving": stats,
        "alone_equals_mixed": equal,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--envs", type=int, default=None)
    ap.add_argument("--workload", default="franka", choices=["franka", "boxpile", "quadruped", "serving"])
    ap.add_argument("--requests", type=int, default=24, help="serving: number of Poisson requests")
    ap.add_argument("--rate", type=float, default=100.0, help="serving: request arrival rate (1/s)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-drift", action="store_true", help="skip the FP32-vs-FP64 drift sample")
    ap.add_argument("--like-for-like", action=argparse.BooleanOptionalAction, default=True,
                    help="franka: also run the reference's speed benchmark (B=30000, fixed target) via the Scene API")
    ap.add_argument("--fused-config", default=None, metavar="EPB,LDS,SCHED,OCC[,GJK_LDS]",
                    help="fixed fused launch shape (e.g. 1,1,0,0) instead of autotuning")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; it measures one GPU (environments are independent, so "
              "N GPUs would each step their own slab with no per-step exchange)", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))
    torch.cuda.set_device(0)
    if args.workload == "serving":
        serving_bench(args)
        return
    wl = Workload(args.workload)
    B = args.envs or wl.B
    steps = args.steps or wl.steps
    model = wl.model
    dm = rt.DeviceModel(model)
    n_blocks = max(1, (steps + args.warmup + BLOCK - 1) // BLOCK)
    targets = None
    if wl.targets(1, 0) is not None:
        targets = [torch.as_tensor(wl.targets(B, k).T, dtype=torch.float32, device="cuda").contiguous()
                   for k in range(n_blocks)]
    print(f"[bench] {args.workload}: D={model.n_dofs} G={model.n_geoms} pairs={model.n_pairs}; B={B} steps={steps}",
          file=sys.stderr, flush=True)
    tune = None
    if args.fused_config:
        epb, lds, sched, occ, *gjk = (int(x) for x in args.fused_config.split(","))
        dm.set_fused_config(epb, bool(lds), bool(sched), occ, bool(gjk[0]) if gjk else False)
    elif not args.no_autotune:
        probe = make_slab(wl, dm, B)
        probe.step(20, rt.RS_MODE_FUSED)  # tune on a state with contacts, not the initial one
        tune = rt.autotune_fused(probe, steps=10)
        del probe
        print(f"[bench] autotune: {tune}", file=sys.stderr, flush=True)
    wall_f, gpu_ms_f, lat_f, nb_f, slab, timed, gblock = run_path(wl, dm, B, steps, args.warmup, rt.RS_MODE_FUSED,
                                                                  targets, use_graph=True)
    errno = slab.errno()
    print(f"[bench] fused: {wall_f:.3f}s for {steps} steps", file=sys.stderr, flush=True)
    plain = None
    if not args.no_plain:
        # the per-stage path on its own layout ([row][env] planes, coalesced for one lane per environment)
        dm_plain = rt.DeviceModel(model, rt.LAYOUT_PLANES)
        wall_p, gpu_ms_p, lat_p, nb_p, *_ = run_path(wl, dm_plain, B, steps, args.warmup, rt.RS_MODE_PLAIN, targets,
                                                     use_graph=False, n_lat=50)
        plain = {"value": B * steps / wall_p, "ms_per_step": 1e3 * wall_p / steps,
                 "p50_ms": float(np.percentile(lat_p, 50)), "p99_ms": float(np.percentile(lat_p, 99))}
        print(f"[bench] plain: {wall_p:.3f}s", file=sys.stderr, flush=True)
    value = B * steps / wall_f
    bytes_per_env_step = algorithmic_bytes_per_env_step(model, nb_f)
    kernel_ms = gpu_ms_f / steps  # one fused launch per substep, events on the launch stream
    achieved = bytes_per_env_step * B / (kernel_ms * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic("k_fused_wave", wl, B)
    out = {
        "metric": "env-steps/s at Franka×4096 (fused vs plain GPU, FP32); p50/p99 step latency",
        "value": value,
        "unit": "env-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall_f / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl.desc.format(B=B), "n_envs": B,
                   "impl": f"fused substep kernel (one wavefront per env), {gblock}-substep hipGraph replayed "
                           f"{timed['graph_launches']}x in the timed region, {timed['eager_substeps']} eager substeps",
                   "fused_config": dm.fused_config},
        "p50_step_ms": float(np.percentile(lat_f, 50)),
        "p99_step_ms": float(np.percentile(lat_f, 99)),
        "plain": plain,
        "fused_over_plain": (value / plain["value"]) if plain else None,
        "autotune": tune,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic, "traffic_source": traffic_src,
                     "bytes_per_env_step": bytes_per_env_step, "mean_overlap_pairs": nb_f,
                     "kernel": "k_fused_wave", "kernel_ms": kernel_ms},
        "valu_roofline": valu_roofline(wl, kernel_ms, B),
        "errno": errno,
    }
    if args.workload != "franka":
        out["metric"] = f"env-steps/s, {args.workload} (fused vs plain GPU, FP32); p50/p99 step latency"
    try:
        out["hbm_copy_GBps"] = hbm_copy_peak_gbps()
    except Exception as e:  # noqa: BLE001
        out["hbm_copy_GBps"] = None
        print(f"[bench] copy peak failed: {e}", file=sys.stderr)
    if not args.no_cpu_baseline and rank == 0:
        v1, vn, nt, sample = cpu_baseline(wl)
        out["cpu_baseline"] = {"value": v1, "unit": "env-steps/s", "cores": 1, "kind": "port",
                               "sample": f"{sample}, same seeds, FP32 C++ restatement of the step",
                               "value_all_cores": vn, "cores_all": nt, "cpu": cpu_model_name()}
    if not args.no_drift and rank == 0:
        print("[bench] drift vs the FP64 restatement", file=sys.stderr, flush=True)
        out["drift_vs_fp64"] = drift_vs_fp64(wl, dm)
    if args.like_for_like and args.workload == "franka" and rank == 0:
        print("[bench] like-for-like: examples/speed_benchmark/franka.py", file=sys.stderr, flush=True)
        out["like_for_like"] = like_for_like()
    if rank == 0:
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
