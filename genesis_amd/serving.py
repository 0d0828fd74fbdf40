"""Continuous-batching rollout server on one GPU.

Requests (scene type, number of environments, horizon) arrive over time. A request is admitted as soon as its state
slab fits in the HBM budget and a segment slot is free; its state stays resident on the device until its horizon is
reached. Every tick advances every admitted request by one substep in a single shared fused launch
(``runtime.SharedLaunch`` -> ``rs_batch_step``), so requests of different scenes and sizes overlap on the device; a
finished request releases its memory and slot immediately, which admits the next waiting ones. Each environment's
trajectory is the same as when its slab is stepped alone (``tests/test_gpu_serving.py`` checks this bit for bit).

Preemption and request checkpoints (after ``scene.py:1402-1446``, the reference's whole-scene save/load, here per
request): a waiting request of higher ``priority`` that does not fit evicts active requests of lower priority (lowest
first, the latest admitted first among equals), but only when evicting them makes it fit. An evicted request's state
slab is copied whole into pinned host memory and its device memory released; it waits again, and on re-admission
its slab is restored from that copy and continues from the substep it reached, bit for bit. ``save_request`` writes
a request (its description, progress and, once started, its state) to a file from which ``load_request`` rebuilds
it, so a request can also move between servers or processes. ``run(..., stop_after_ticks=N, checkpoint_dir=d)``
stops a run after N shared launches and writes every unfinished request (active, preempted or still waiting) to d;
``load_checkpoint(d)`` returns them, and a later ``run`` of any server finishes them bit for bit as if uninterrupted.
"""

from __future__ import annotations

import ctypes
import json
import os
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
    priority: int = 0
    preemptions: int = 0
    host_state: torch.Tensor | None = None  # the slab of a started request that is not resident (pinned host memory)


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

    def __init__(self, hbm_budget_bytes: int, max_active: int = 32, stream: torch.cuda.Stream | None = None,
                 preempt: bool = True):
        self.budget = int(hbm_budget_bytes)
        self.max_active = max_active
        self.preempt = preempt
        self.preemptions = 0
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

    # ------------------------------------------------------------------ residency
    def admit(self, req: Request) -> None:
        """Give a request its device slab: restored from its host copy if it was started, else its initial state."""
        if req.slab is not None:
            raise ValueError(f"request {req.rid} is already resident")
        if not req.nbytes:
            req.nbytes = self.slab_bytes(req.scene, req.n_envs)
        with torch.cuda.stream(self.stream):
            if req.host_state is not None:
                slab = rt.StateSlab(self.model(req.scene), req.n_envs)
                if slab.nbytes != req.host_state.numel():
                    raise ValueError(f"request {req.rid}: saved state holds {req.host_state.numel()} bytes, its slab "
                                     f"{slab.nbytes}")
                slab.buffer.copy_(req.host_state)
                req.host_state = None
            else:
                slab = self.new_slab(req)
        req.slab = slab
        self.used += req.nbytes

    def evict(self, req: Request) -> None:
        """Copy a resident request's slab whole into pinned host memory and release its device memory."""
        if req.slab is None:
            raise ValueError(f"request {req.rid} is not resident")
        with torch.cuda.stream(self.stream):
            host = torch.empty(req.slab.nbytes, dtype=torch.uint8, pin_memory=True)
            host.copy_(req.slab.buffer)  # synchronous: the copy has landed when it returns
        req.host_state = host
        req.slab = None
        req.preemptions += 1
        self.preemptions += 1
        self.used -= req.nbytes

    def tick(self, active: list[Request]) -> float:
        """Advance every request in ``active`` by one substep in one shared launch; the launch's time in ms."""
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(self.stream):
            e0.record(self.stream)
            self.shared.step([r.slab for r in active], self.stream)
            e1.record(self.stream)
        e1.synchronize()
        for r in active:
            r.steps_done += 1
        return e0.elapsed_time(e1)

    # ------------------------------------------------------------------ checkpoints
    def save_request(self, req: Request, path) -> None:
        """A request's description, progress and (once started) full state to one ``.npz`` file."""
        meta = np.array([req.rid, req.n_envs, req.horizon, req.steps_done, req.priority], dtype=np.int64)
        arrays = {"meta": meta, "scene": np.array(req.scene), "arrival": np.array(req.arrival, dtype=np.float64)}
        if req.slab is not None:
            torch.cuda.current_stream().wait_stream(self.stream)
            arrays["state"] = req.slab.buffer.cpu().numpy()
        elif req.host_state is not None:
            arrays["state"] = req.host_state.numpy()
        with open(path, "wb") as f:
            np.savez(f, **arrays)

    def load_request(self, path) -> Request:
        """Rebuild a request written by :meth:`save_request`; a started one resumes from its saved substep."""
        with np.load(path, allow_pickle=False) as z:
            rid, n_envs, horizon, steps_done, priority = (int(x) for x in z["meta"])
            scene = str(z["scene"])
            if scene not in SCENES:
                raise ValueError(f"{path!s}: unknown scene {scene!r}")
            req = Request(rid, scene, n_envs, horizon, float(z["arrival"]), steps_done=steps_done, priority=priority)
            req.nbytes = self.slab_bytes(scene, n_envs)
            if "state" in z.files:
                state = z["state"]
                if state.dtype != np.uint8 or state.size != req.nbytes:
                    raise ValueError(f"{path!s}: state of {state.size} bytes, a {scene} slab of {n_envs} "
                                     f"environments holds {req.nbytes}")
                host = torch.empty(req.nbytes, dtype=torch.uint8, pin_memory=True)
                host.copy_(torch.from_numpy(state))
                req.host_state = host
            elif steps_done:
                raise ValueError(f"{path!s}: a started request without its state")
        return req

    def _victims(self, head: Request, active: list[Request]) -> list[Request] | None:
        """Active requests of lower priority whose eviction lets ``head`` in (lowest priority first, the latest
        admitted first among equals), or None when evicting all of them would not."""
        cands = sorted((r for r in active if r.priority < head.priority),
                       key=lambda r: (r.priority, -(r.admitted or 0.0)))
        out, used, n = [], self.used, len(active)
        for r in cands:
            if n < self.max_active and used + head.nbytes <= self.budget:
                break
            out.append(r)
            used -= r.nbytes
            n -= 1
        return out if (n < self.max_active and used + head.nbytes <= self.budget) else None

    # ------------------------------------------------------------------ loop
    def save_checkpoint(self, directory, requests: list[Request]) -> None:
        """Every request of ``requests`` to ``directory`` (one ``save_request`` file each, and an index)."""
        os.makedirs(directory, exist_ok=True)
        names = []
        for r in requests:
            name = f"request_{r.rid}.npz"
            self.save_request(r, os.path.join(directory, name))
            names.append(name)
        with open(os.path.join(directory, "index.json"), "w") as f:
            json.dump({"requests": names}, f)

    def load_checkpoint(self, directory) -> list[Request]:
        """The requests a ``save_checkpoint`` wrote, ready for ``run`` (started ones resume from their state)."""
        with open(os.path.join(directory, "index.json")) as f:
            names = json.load(f)["requests"]
        return [self.load_request(os.path.join(directory, n)) for n in names]

    def run(self, requests: list[Request], keep_final: bool = False, stop_after_ticks: int | None = None,
            checkpoint_dir=None) -> dict:
        """Serve ``requests`` until all are finished, or until ``stop_after_ticks`` shared launches when it is given:
        then the unfinished requests are written to ``checkpoint_dir`` (``save_checkpoint``) and released."""
        if stop_after_ticks is not None and checkpoint_dir is None:
            raise ValueError("stop_after_ticks needs a checkpoint_dir")
        waiting = sorted(requests, key=lambda r: r.arrival)
        start = {id(r): r.steps_done for r in waiting}  # requests loaded from a checkpoint resume part-way
        for r in waiting:
            r.nbytes = self.slab_bytes(r.scene, r.n_envs)
            if r.nbytes > self.budget:
                raise ValueError(f"request {r.rid} needs {r.nbytes} bytes, more than the budget {self.budget}")
            if r.steps_done >= r.horizon or r.slab is not None:
                raise ValueError(f"request {r.rid} is finished or already resident")
        active: list[Request] = []
        done: list[Request] = []
        ticks = 0
        stopped = False
        t0 = time.perf_counter()
        with torch.cuda.stream(self.stream):
            while waiting or active:
                if stop_after_ticks is not None and ticks >= stop_after_ticks:
                    self.save_checkpoint(checkpoint_dir, active + waiting)
                    for r in active:
                        r.slab = None
                        self.used -= r.nbytes
                    stopped = True
                    break
                now = time.perf_counter() - t0
                while True:
                    # the arrived request of highest priority, the earliest among equals
                    arrived = [r for r in waiting if r.arrival <= now]
                    if not arrived:
                        break
                    head = min(arrived, key=lambda r: (-r.priority, r.arrival, r.rid))
                    if len(active) >= self.max_active or self.used + head.nbytes > self.budget:
                        victims = self._victims(head, active) if self.preempt else None
                        if not victims:
                            break
                        for v in victims:
                            self.evict(v)
                            active.remove(v)
                            waiting.append(v)
                    waiting.remove(head)
                    self.admit(head)
                    if head.admitted is None:
                        head.admitted = time.perf_counter() - t0
                    active.append(head)
                if not active:
                    time.sleep(max(0.0, min(min(r.arrival for r in waiting) - now, 0.01)))
                    continue
                self.tick_ms.append(self.tick(active))
                self.active_per_tick.append(len(active))
                ticks += 1
                still = []
                for r in active:
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
        env_steps = sum(r.n_envs * (r.horizon - start[id(r)]) for r in done)
        busy = sum(self.tick_ms) * 1e-3
        return {
            "requests": len(done),
            "env_steps": env_steps,
            "wall_s": wall,
            "env_steps_per_s": env_steps / wall,
            "env_steps_per_busy_s": env_steps / busy if busy > 0 else None,
            "ticks": len(self.tick_ms),
            "p50_step_ms": float(np.percentile(self.tick_ms, 50)) if self.tick_ms else None,
            "p99_step_ms": float(np.percentile(self.tick_ms, 99)) if self.tick_ms else None,
            "mean_active": float(np.mean(self.active_per_tick)) if self.active_per_tick else 0.0,
            "max_active": int(np.max(self.active_per_tick)) if self.active_per_tick else 0,
            "mean_queue_s": float(np.mean([r.admitted - r.arrival for r in done])) if done else None,
            "preemptions": self.preemptions,
            "stopped": stopped,
        }
