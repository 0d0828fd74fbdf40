"""Serving loop: requests of different scenes, sizes and horizons co-batched into shared fused launches.

Each request's result must be the same alone and in the mix: the environments are independent and the fused step is
deterministic, so the check is bit-exact. The same holds for a request that is evicted to host memory and restored
(preemption), or saved to a file and rebuilt from it (request checkpoints, after scene.py:1402-1446).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from genesis_amd import runtime as rt
from genesis_amd.serving import Request, RolloutServer, poisson_requests

pytestmark = pytest.mark.gpu


def _alone(server, req):
    slab = server.new_slab(req)
    slab.step(req.horizon, rt.RS_MODE_FUSED)
    torch.cuda.synchronize()
    return slab["qpos"].clone(), slab["vel"].clone()


def test_mixed_requests_equal_alone():
    reqs = [Request(0, "franka", 100, 30, 0.0), Request(1, "quadruped", 64, 45, 0.0),
            Request(2, "boxpile", 70, 25, 0.0), Request(3, "franka", 33, 12, 0.0),
            Request(4, "quadruped", 129, 20, 0.0)]
    server = RolloutServer(hbm_budget_bytes=64 << 30, max_active=8)
    print("[serving] mixed run", flush=True)
    stats = server.run(reqs, keep_final=True)
    assert stats["requests"] == len(reqs)
    assert stats["max_active"] == len(reqs)
    for r in reqs:
        print(f"[serving] replay request {r.rid} ({r.scene}, {r.n_envs} envs, {r.horizon} steps)", flush=True)
        q, v = _alone(server, r)
        assert torch.equal(q, r.final_qpos), r.rid
        assert torch.equal(v, r.final_vel), r.rid


def test_budget_admission_queues_and_completes():
    reqs = poisson_requests(8, rate_hz=200.0, seed=4, min_envs=32, max_envs=256, min_horizon=10, max_horizon=40)
    probe = RolloutServer(hbm_budget_bytes=1 << 40)
    need = max(probe.slab_bytes(r.scene, r.n_envs) for r in reqs)
    server = RolloutServer(hbm_budget_bytes=2 * need, max_active=8)  # at most two of the largest at a time
    stats = server.run(reqs)
    assert stats["requests"] == len(reqs)
    assert server.used == 0
    assert stats["p99_step_ms"] >= stats["p50_step_ms"] > 0


def test_serving_mix_at_its_stated_size():
    """The serving-mix workload as stated: 24 seeded Poisson requests at 100/s, scenes uniform over franka / box pile /
    quadruped, n_envs log-uniform in [64, 4096], horizons uniform in [50, 500]. Every request ends bit-identical to its
    run alone; and each request's alone run follows the FP32 restatement over its first 5 substeps (per-environment
    max |q_device - q_restatement|, 99% of the environments within 1e-3, every environment within 5e-2: contact-rich
    states amplify FP32 rounding differences)."""
    from tests.oracle_lib import Oracle
    from tests.parity import get_dev, slab_to_oracle

    reqs = poisson_requests(24, rate_hz=100.0, seed=4, min_envs=64, max_envs=4096, min_horizon=50, max_horizon=500)
    server = RolloutServer(hbm_budget_bytes=64 << 30, max_active=32)
    print(f"[serving] stated mix: {sum(r.n_envs * r.horizon for r in reqs)} env-steps", flush=True)
    stats = server.run(reqs, keep_final=True)
    assert stats["requests"] == 24 and server.used == 0
    print(f"[serving] {stats['env_steps_per_s']:.3e} env-steps/s, mean active {stats['mean_active']:.1f}, "
          f"tick p50 {stats['p50_step_ms']:.2f} ms p99 {stats['p99_step_ms']:.2f} ms", flush=True)
    worst = []
    for r in reqs:
        q, v = _alone(server, r)
        assert torch.equal(q, r.final_qpos) and torch.equal(v, r.final_vel), f"request {r.rid} differs from its run alone"
        slab = server.new_slab(r)
        oracle = Oracle(slab.dmodel.model, r.n_envs, precision=32)
        slab_to_oracle(slab, oracle)
        slab.step(5, rt.RS_MODE_FUSED)
        torch.cuda.synchronize()
        oracle.step(5, 16)
        d = np.abs(get_dev(slab, "qpos") - oracle.get("qpos")).max(axis=1)
        worst.append(float(d.max()))
        print(f"[serving] request {r.rid} ({r.scene}, {r.n_envs} envs, {r.horizon} steps): alone == mixed; 5 steps vs "
              f"restatement: p99 {np.percentile(d, 99):.2e} max {d.max():.2e}", flush=True)
        assert np.percentile(d, 99) <= 1e-3 and d.max() <= 5e-2, r.rid


def _alone_like(server, req):
    return _alone(server, Request(req.rid, req.scene, req.n_envs, req.horizon, 0.0))


def test_preemption_evicts_lower_priority_and_resumes_bit_exact():
    """Two long requests of priority 0 fill the budget; a priority-1 request arriving later does not fit, evicts the
    later-admitted of them to pinned host memory, runs, and the evicted one resumes when memory frees. Every request
    ends bit-equal to its run alone."""
    lo = [Request(0, "franka", 256, 600, 0.0), Request(1, "quadruped", 256, 600, 0.0)]
    hi = Request(2, "boxpile", 64, 40, 0.05, priority=1)
    probe = RolloutServer(hbm_budget_bytes=1 << 40)
    need = [probe.slab_bytes(r.scene, r.n_envs) for r in lo + [hi]]
    server = RolloutServer(hbm_budget_bytes=sum(need) - 1, max_active=8)  # the three never fit together
    print(f"[serving] preemption run, slabs {need} bytes", flush=True)
    stats = server.run(lo + [hi], keep_final=True)
    print(f"[serving] preemptions {stats['preemptions']}, per request {[r.preemptions for r in lo + [hi]]}", flush=True)
    assert stats["requests"] == 3 and server.used == 0
    assert stats["preemptions"] >= 1 and hi.preemptions == 0 and lo[0].preemptions + lo[1].preemptions >= 1
    assert hi.finished < max(lo[0].finished, lo[1].finished)
    for r in lo + [hi]:
        q, v = _alone_like(server, r)
        assert torch.equal(q, r.final_qpos) and torch.equal(v, r.final_vel), f"request {r.rid} differs from alone"


def test_request_eviction_and_file_checkpoint_round_trip(tmp_path):
    """A request stepped 25 substeps, evicted and restored, stepped 10 more, saved to a file and rebuilt from it as a
    new request object, then finished: bit-equal to its run alone. A file for another slab size is refused."""
    import os

    server = RolloutServer(hbm_budget_bytes=64 << 30)
    r = Request(7, "quadruped", 128, 60, 0.0)
    server.admit(r)
    for _ in range(25):
        server.tick([r])
    server.evict(r)
    assert r.slab is None and r.host_state is not None and server.used == 0
    server.admit(r)
    for _ in range(10):
        server.tick([r])
    path = os.path.join(str(tmp_path), "req7.npz")
    server.save_request(r, path)
    r2 = server.load_request(path)
    assert (r2.rid, r2.scene, r2.n_envs, r2.horizon, r2.steps_done) == (7, "quadruped", 128, 60, 35)
    server.admit(r2)
    while r2.steps_done < r2.horizon:
        server.tick([r2])
    q, v = _alone_like(server, r)
    assert torch.equal(q, r2.slab["qpos"]) and torch.equal(v, r2.slab["vel"])
    # a request that has not started round-trips without state; a state of the wrong size is refused
    fresh = Request(8, "franka", 32, 10, 0.0)
    p2 = os.path.join(str(tmp_path), "req8.npz")
    server.save_request(fresh, p2)
    back = server.load_request(p2)
    assert back.host_state is None and back.steps_done == 0
    with np.load(path, allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    arrays["meta"] = arrays["meta"].copy()
    arrays["meta"][1] = 64  # 64 environments: another slab size
    p3 = os.path.join(str(tmp_path), "bad.npz")
    with open(p3, "wb") as f:
        np.savez(f, **arrays)
    with pytest.raises(ValueError):
        server.load_request(p3)


def test_stopped_run_resumes_from_its_checkpoint_bit_exact(tmp_path):
    """A run stopped after 25 shared launches writes its unfinished requests (active ones with their state, one not
    yet admitted without) to a checkpoint; a fresh server loads it and finishes them: every request ends bit-equal to
    its run alone."""
    import os

    ck = os.path.join(str(tmp_path), "ck")
    reqs = [Request(0, "franka", 64, 80, 0.0), Request(1, "quadruped", 96, 60, 0.0), Request(2, "boxpile", 32, 40, 0.0),
            Request(3, "franka", 48, 30, 30.0)]  # arrives long after the stop
    first = RolloutServer(hbm_budget_bytes=64 << 30, max_active=8)
    stats = first.run(reqs, stop_after_ticks=25, checkpoint_dir=ck)
    assert stats["stopped"] and stats["requests"] == 0 and first.used == 0
    second = RolloutServer(hbm_budget_bytes=64 << 30, max_active=8)
    back = second.load_checkpoint(ck)
    assert sorted((r.rid, r.steps_done, r.host_state is not None) for r in back) == [
        (0, 25, True), (1, 25, True), (2, 25, True), (3, 0, False)]
    for r in back:
        r.arrival = 0.0
    print("[serving] resuming 4 requests from the checkpoint", flush=True)
    stats = second.run(back, keep_final=True)
    assert stats["requests"] == 4 and not stats["stopped"] and second.used == 0
    for r in back:
        q, v = _alone_like(second, r)
        assert torch.equal(q, r.final_qpos) and torch.equal(v, r.final_vel), f"request {r.rid} differs from alone"
