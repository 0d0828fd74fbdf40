"""Serving loop: requests of different scenes, sizes and horizons co-batched into shared fused launches.

Each request's result must be the same alone and in the mix: the environments are independent and the fused step is
deterministic, so the check is bit-exact.
"""

from __future__ import annotations

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
