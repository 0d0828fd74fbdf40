"""Continuous-batching rollout server on one GPU.

Requests (scene type, number of environments, horizon) arrive over time. A request is admitted as soon as its state
slab fits in the HBM budget and a segment slot is free; its state stays resident on the device until its horizon is
reached. Every tick advances every admitted request by one substep in a single shared fused launch
(``runtime.SharedLaunch`` -> ``rs_batch_step``), so requests of different scenes and sizes overlap on the device; a
finished request releases its memory and slot immediately, which admits the next waiting ones. Each environment's
trajectory is the same as when its slab is stepped alone (``tests/test_gpu_serving.py`` checks this bit for bit).
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import runtime as rt
from . import scenes

SCENES = ("franka", "boxpile", "quadruped")


@dataclass
class Request:
    rid: int
    scene: str
    n_envs: int
    horizon: int
    arrival: float  # seconds after the start of the run
    admitted: float | None = None
    finished: float | None = None
    steps_done: int = 0
    slab: rt.StateSlab | None = None
    nbytes: int = 0
    final_qpos: torch.Tensor | None = None
    final_vel: torch.Tensor | None = None


def poisson_requests(n: int, rate_hz: float, seed: int = 4, min_envs: int = 64, max_envs: int = 4096,
                     min_horizon: int = 50, max_horizon: int = 500) -> list[Request]:
    """Seeded Poisson arrivals; scene uniform over SCENES, n_envs log-uniform in [min_envs, max_envs], horizon uniform
    in [min_horizon, max_horizon]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = np.cumsum(rng.exponential(1.0 / rate_hz, n))
    out = []
    for i in range(n):
        scene = SCENES[int(rng.integers(len(SCENES)))]
        envs = int(round(float(np.exp(rng.uniform(np.log(min_envs), np.log(max_envs))))))
        horizon = int(rng.integers(min_horizon, max_horizon + 1))
        out.append(Request(i, scene, envs, horizon, float(t[i])))
    return out


class RolloutServer:
    """Admission by HBM budget, shared fused launches, per-tick latency accounting."""

    def __init__(self, hbm_budget_bytes: int, max_active: int = 32, stream: torch.cuda.Stream | None = None):
        self.budget = int(hbm_budget_bytes)
        self.max_active = max_active
        self.stream = stream or torch.cuda.Stream()
        self.shared = rt.SharedLaunch(max_active)
        self._models: dict[str, rt.DeviceModel] = {}
        self.used = 0
        self.tick_ms: list[float] = []
        self.active_per_tick: list[int] = []

    # ------------------------------------------------------------------ scenes
    def model(self, scene: str) -> rt.DeviceModel:
        if scene not in self._models:
            m = {"franka": scenes.franka_model, "boxpile": scenes.box_pile_model,
                 "quadruped": scenes.quadruped_model}[scene]()
            self._models[scene] = rt.DeviceModel(m)
        return self._models[scene]

    def slab_bytes(self, scene: str, n_envs: int) -> int:
        dm = self.model(scene)
        total = ctypes.c_int64()
        rt._check(rt.load_library().rs_state_layout(dm.handle, n_envs, None, ctypes.byref(total)))
        return total.value

    @staticmethod
    def init_state(req: Request, slab: rt.StateSlab):
        """Seeded initial state of a request (seeded by its id, so that a replay reproduces it)."""
        m = slab.dmodel.model
        B = slab.n_envs
        if req.scene == "franka":
            slab["qpos"].copy_(torch.as_tensor(scenes.franka_initial_qpos(m, B, seed=req.rid).T, dtype=torch.float32))
            slab["ctrl_mode"].fill_(0)
            slab["ctrl_pos"].copy_(torch.as_tensor(scenes.franka_targets(m, B, req.rid).T, dtype=torch.float32))
        elif req.scene == "quadruped":
            slab["qpos"].copy_(torch.as_tensor(scenes.quadruped_initial_qpos(m, B).T, dtype=torch.float32))
            slab["ctrl_mode"].copy_(torch.as_tensor(scenes.quadruped_ctrl_mode(m, B).T, dtype=torch.int32))
            slab["ctrl_pos"].copy_(torch.as_tensor(scenes.quadruped_targets(m, B, req.rid).T, dtype=torch.float32))
        slab.forward_kinematics()

    def new_slab(self, req: Request) -> rt.StateSlab:
        slab = rt.StateSlab(self.model(req.scene), req.n_envs)
        self.init_state(req, slab)
        return slab

    # ------------------------------------------------------------------ loop
    def run(self, requests: list[Request], keep_final: bool = False) -> dict:
        waiting = sorted(requests, key=lambda r: r.arrival)
        for r in waiting:
            r.nbytes = self.slab_bytes(r.scene, r.n_envs)
            if r.nbytes > self.budget:
                raise ValueError(f"request {r.rid} needs {r.nbytes} bytes, more than the budget {self.budget}")
        active: list[Request] = []
        done: list[Request] = []
        t0 = time.perf_counter()
        with torch.cuda.stream(self.stream):
            while waiting or active:
                now = time.perf_counter() - t0
                while waiting and waiting[0].arrival <= now and len(active) < self.max_active \
                        and self.used + waiting[0].nbytes <= self.budget:
                    r = waiting.pop(0)
                    r.slab = self.new_slab(r)
                    r.admitted = time.perf_counter() - t0
                    self.used += r.nbytes
                    active.append(r)
                if not active:
                    time.sleep(max(0.0, min(waiting[0].arrival - now, 0.01)))
                    continue
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(self.stream)
                self.shared.step([r.slab for r in active], self.stream)
                e1.record(self.stream)
                e1.synchronize()
                self.tick_ms.append(e0.elapsed_time(e1))
                self.active_per_tick.append(len(active))
                still = []
                for r in active:
                    r.steps_done += 1
                    if r.steps_done >= r.horizon:
                        r.finished = time.perf_counter() - t0
                        if keep_final:
                            r.final_qpos = r.slab["qpos"].clone()
                            r.final_vel = r.slab["vel"].clone()
                        r.slab = None
                        self.used -= r.nbytes
                        done.append(r)
                    else:
                        still.append(r)
                active = still
        wall = time.perf_counter() - t0
        env_steps = sum(r.n_envs * r.horizon for r in done)
        busy = sum(self.tick_ms) * 1e-3
        return {
            "requests": len(done),
            "env_steps": env_steps,
            "wall_s": wall,
            "env_steps_per_s": env_steps / wall,
            "env_steps_per_busy_s": env_steps / busy if busy > 0 else None,
            "ticks": len(self.tick_ms),
            "p50_step_ms": float(np.percentile(self.tick_ms, 50)),
            "p99_step_ms": float(np.percentile(self.tick_ms, 99)),
            "mean_active": float(np.mean(self.active_per_tick)),
            "max_active": int(np.max(self.active_per_tick)),
            "mean_queue_s": float(np.mean([r.admitted - r.arrival for r in done])),
        }
