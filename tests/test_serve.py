"""The serving scheduler's queue and slot policy (no GPU): admission only
after arrival and only into free slots, lowest slot first, slots freed
when a request's last step has run, latency percentiles."""

from __future__ import annotations

import numpy as np
import pytest

from flybrain_amd.serve import (Scheduler, SimRequest, SimResult,
                                bytes_per_trial, check_request,
                                poisson_requests, summarize)


def req(rid, steps, arrival=0.0):
    return SimRequest(rid=rid, steps=steps, drive=np.array([1, 2]),
                      rate_hz=100.0, arrival_s=arrival)


def test_admission_waits_for_arrival_and_free_slots():
    s = Scheduler(2)
    for k, (steps, arr) in enumerate([(40, 0.0), (20, 0.0), (60, 0.5),
                                      (10, 0.1)]):
        s.submit(req(k, steps, arr))
    got = s.admit(now_s=0.05, t=0)
    assert [(slot, r.rid) for slot, r, _ in got] == [(0, 0), (1, 1)]
    assert s.admit(now_s=1.0, t=0) == []          # no free slot
    assert s.finished(t=19) == []
    done = s.finished(t=20)                        # request 1 ran 20 steps
    assert [(slot, r.rid, t0) for slot, r, t0, _ in done] == [(1, 1, 0)]
    got = s.admit(now_s=0.2, t=20)  # request 3 arrived at 0.1, 2 not yet
    assert [(slot, r.rid) for slot, r, _ in got] == [(1, 3)]
    assert s.finished(t=29) == [] and len(s.finished(t=30)) == 1
    got = s.admit(now_s=0.6, t=40)
    assert [(slot, r.rid) for slot, r, _ in got] == [(1, 2)]


def test_fifo_order_and_slot_reuse():
    s = Scheduler(3)
    for k in range(7):
        s.submit(req(k, 10 * (k + 1)))
    t, order, slots_seen = 0, [], set()
    while s.queue or s.busy:
        for slot, r, _ in s.admit(now_s=0.0, t=t):
            order.append(r.rid)
            slots_seen.add(slot)
        t += 10
        for slot, r, t0, _ in s.finished(t):
            assert t0 + r.steps <= t < t0 + r.steps + 10
    assert order == list(range(7))
    assert slots_seen == {0, 1, 2}


def test_rejects_empty_request():
    with pytest.raises(ValueError):
        Scheduler(1).submit(req(0, 0))
    with pytest.raises(ValueError):
        Scheduler(0)


def test_summary_percentiles():
    res = [SimResult(k, np.zeros(3, np.int64), 1000, 0.0, 0.0,
                     0.001 * (k + 1), 0) for k in range(100)]
    st = summarize(res, wall=2.0, slots=4, windows=10, window=20)
    assert st["requests"] == 100
    assert st["p50_ms"] == pytest.approx(50.5)
    assert st["p99_ms"] == pytest.approx(99.01)
    assert st["trial_s_per_s"] == pytest.approx(100 * 1000 * 1e-4 / 2.0)


def test_workload_generator_and_memory_model():
    rs = poisson_requests(1000, 50, rate_req_s=100.0, seed=3)
    arr = np.array([r.arrival_s for r in rs])
    assert np.all(np.diff(arr) > 0)
    assert 0.2 < arr[-1] < 1.0                    # ~50 / 100 req/s
    assert all(500 <= r.steps <= 20_000 for r in rs)
    assert all(20 <= r.drive.size <= 200 for r in rs)
    assert all(50.0 <= r.rate_hz <= 250.0 for r in rs)
    assert all(r.silenced.size <= 50 for r in rs)
    assert bytes_per_trial(138_639, "pull") == 193 * 138_639


def test_requests_are_checked_before_they_reach_a_slot():
    """Bad requests are named at submission (steps, indices, drive size,
    rate, seed); a good one passes."""
    n = 100
    good = SimRequest(rid=0, steps=40, drive=np.array([1, 2, 99]),
                      rate_hz=150.0, silenced=np.array([5]), seed=3)
    assert check_request(good, n, max_drive=8) is None
    bad = {
        "steps": dict(steps=0),
        "steps ": dict(steps=2.5),
        "drive index": dict(drive=np.array([1, 100])),
        "drive index ": dict(drive=np.array([-1])),
        "silenced index": dict(silenced=np.array([100])),
        "indices": dict(drive=np.array([1.0, 2.0])),
        "max_drive": dict(drive=np.arange(9)),
        "rate": dict(rate_hz=10_001.0),
        "rate ": dict(rate_hz=-1.0),
        "rate  ": dict(rate_hz=float("nan")),
        "seed": dict(seed=-5),
    }
    for what, kw in bad.items():
        r = SimRequest(**{**vars(good), **kw})
        why = check_request(r, n, max_drive=8)
        assert why is not None, what
    # a repeated drive index counts once against max_drive
    r = SimRequest(**{**vars(good), "drive": np.array([1] * 20)})
    assert check_request(r, n, max_drive=1) is None


def test_release_drop_and_resume_order():
    """A released slot is reused; a queued request can be dropped; a
    resumed request is admitted before new arrivals and keeps its own
    clock (t_start = t - steps done)."""
    s = Scheduler(1)
    for k in range(3):
        s.submit(req(k, 100))
    (slot, r, done), = s.admit(now_s=0.0, t=0)
    assert (slot, r.rid, done) == (0, 0, 0)
    assert s.slot_of(0) == 0 and s.slot_of(1) is None
    assert s.drop_queued(1) and not s.drop_queued(1)
    got = s.release(0)
    assert got[0].rid == 0 and s.free == [0]
    s.resumed.append((got[0], 40, 0.0))
    (slot, r, done), = s.admit(now_s=0.0, t=60)
    assert (r.rid, done) == (0, 40) and s.active[slot][1] == 20
    assert s.finished(t=119) == [] and len(s.finished(t=120)) == 1
    (slot, r, done), = s.admit(now_s=0.0, t=120)
    assert (r.rid, done) == (2, 0)


def test_summary_leaves_out_rejected_and_cancelled():
    ok = [SimResult(k, np.zeros(3, np.int64), 100, 0.0, 0.0, 0.01, 0)
          for k in range(4)]
    rej = SimResult(9, None, 100, 0.0, 0.0, 0.0, -1, error="bad")
    can = SimResult(8, np.zeros(3, np.int64), 100, 0.0, 0.0, 5.0, 0,
                    steps_done=40, cancelled=True)
    assert rej.steps_done == 0 and not rej.ok and not can.ok
    st = summarize(ok + [rej, can], wall=1.0, slots=2, windows=5, window=20)
    assert st["requests"] == 4 and st["rejected"] == 1
    assert st["cancelled"] == 1
    assert st["p99_ms"] == pytest.approx(10.0)
    assert st["trial_s_per_s"] == pytest.approx((400 + 40) * 1e-4)
