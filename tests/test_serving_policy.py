"""The serving loop's preemption policy (genesis_amd/serving.py RolloutServer._victims) on the CPU: which active
requests a waiting one evicts, and when it evicts none. No device is touched (the server object is built without its
stream and shared launch)."""

from __future__ import annotations

from genesis_amd.serving import Request, RolloutServer


def _server(budget, max_active=8, used=0):
    s = RolloutServer.__new__(RolloutServer)  # the policy needs only these fields
    s.budget, s.max_active, s.used = budget, max_active, used
    return s


def _req(rid, nbytes, priority=0, admitted=0.0):
    r = Request(rid, "franka", 1, 1, 0.0, priority=priority)
    r.nbytes, r.admitted = nbytes, admitted
    return r


def test_evicts_lowest_priority_latest_admitted_first_and_only_what_is_needed():
    a, b, c = _req(0, 40, priority=0, admitted=0.1), _req(1, 40, priority=0, admitted=0.2), _req(2, 20, priority=1)
    s = _server(budget=100, used=100)
    head = _req(3, 30, priority=2)
    # lowest priority first (a and b, 0), among them the later admitted (b); b frees 40 >= the 30 needed
    assert s._victims(head, [a, b, c]) == [b]
    head2 = _req(4, 70, priority=2)
    assert s._victims(head2, [a, b, c]) == [b, a]


def test_never_evicts_equal_or_higher_priority_and_gives_up_when_eviction_cannot_make_room():
    a, b = _req(0, 50, priority=1), _req(1, 50, priority=3)
    s = _server(budget=100, used=100)
    assert s._victims(_req(2, 10, priority=1), [a, b]) is None  # equal priority: no preemption
    assert s._victims(_req(3, 60, priority=2), [a, b]) is None  # evicting a (50) is not enough for 60
    assert s._victims(_req(4, 50, priority=2), [a, b]) == [a]


def test_slot_limit_counts_like_memory():
    a, b = _req(0, 1, priority=0, admitted=0.1), _req(1, 1, priority=0, admitted=0.2)
    s = _server(budget=1000, max_active=2, used=2)
    assert s._victims(_req(2, 1, priority=1), [a, b]) == [b]
