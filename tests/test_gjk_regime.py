"""How often each workload meets the MPR -> GJK/EPA hand-over (narrowphase.py:726-734).

With the default MPR collision path the reference re-runs a detection with GJK/EPA when MPR reports a penetration
deeper than the pair's tolerance and either no contact direction was cached for the pair or the overlap is large
(mc_tolerance * penetration >= mpr_to_gjk_overlap_ratio * tolerance). Both the CPU restatement and the device kernels
take the hand-over; the restatement counts those detections (n_gjk_regime: first detections,
n_gjk_regime_perturbed: the multi-contact perturbed ones), which tells how much of each workload's parity rests on the
GJK/EPA path:

- the legged workload (sphere feet and capsule legs on a plane, analytic contacts) never reaches it;
- the box pile reaches it only on impact frames, for a small fraction of its contacts;
- the Franka workload with joint angles drawn uniformly within the limits reaches it for a large share of its
  contacts (deep self-penetration between hull geoms); its device contact sets are compared with the restatement's
  in the GPU stage-parity tests.
"""

from __future__ import annotations

import numpy as np

from genesis_amd.scenes import (box_pile_model, franka_initial_qpos, franka_model, franka_targets,
                                quadruped_ctrl_mode, quadruped_initial_qpos, quadruped_model, quadruped_targets)
from tests.oracle_lib import Oracle


def _count(o, steps, retarget=None):
    n0 = n1 = contacts = 0
    for k in range(steps):
        if retarget is not None and k % 100 == 0:
            retarget(k // 100)
        o.step(1, 8)
        n0 += int(o.get("n_gjk_regime").sum())
        n1 += int(o.get("n_gjk_regime_perturbed").sum())
        contacts += int(o.get("n_contacts").sum())
    return n0, n1, contacts


def test_quadruped_never_reaches_the_gjk_hand_over():
    model, B = quadruped_model(), 16
    o = Oracle(model, B, precision=32)
    o.set("qpos", quadruped_initial_qpos(model, B))
    o.set("ctrl_mode", quadruped_ctrl_mode(model, B))
    o.set("ctrl_pos", quadruped_targets(model, B, 0))
    o.forward_kinematics()
    n0, n1, contacts = _count(o, 200)
    print(f"quadruped B={B} 200 steps: contacts {contacts}, GJK regime {n0} + {n1} perturbed", flush=True)
    assert contacts > 0 and n0 == 0 and n1 == 0


def test_box_pile_reaches_it_on_impact_frames_only():
    model, B = box_pile_model(), 4
    o = Oracle(model, B, precision=32)
    o.forward_kinematics()
    n0, n1, contacts = _count(o, 300)
    print(f"box pile B={B} 300 steps: contacts {contacts}, GJK regime {n0} + {n1} perturbed", flush=True)
    assert 0 < n0 < 0.02 * contacts


def test_franka_random_configurations_reach_it_often():
    model, B = franka_model(), 16
    o = Oracle(model, B, precision=32)
    o.set("qpos", franka_initial_qpos(model, B, 0))
    o.set("ctrl_mode", np.zeros((B, model.n_dofs)))
    o.forward_kinematics()
    n0, n1, contacts = _count(o, 200, lambda k: o.set("ctrl_pos", franka_targets(model, B, k)))
    print(f"franka B={B} 200 steps: contacts {contacts}, GJK regime {n0} + {n1} perturbed", flush=True)
    assert n0 > 0.05 * contacts
