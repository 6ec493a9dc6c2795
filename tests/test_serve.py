"""The serving scheduler's queue and slot policy (no GPU): admission only
after arrival and only into free slots, lowest slot first, slots freed
when a request's last step has run, latency percentiles."""

from __future__ import annotations

import numpy as np
import pytest

from flybrain_amd.serve import (Scheduler, SimRequest, SimResult,
                                bytes_per_trial, poisson_requests, summarize)


def req(rid, steps, arrival=0.0):
    return SimRequest(rid=rid, steps=steps, drive=np.array([1, 2]),
                      rate_hz=100.0, arrival_s=arrival)


def test_admission_waits_for_arrival_and_free_slots():
    s = Scheduler(2)
    for k, (steps, arr) in enumerate([(40, 0.0), (20, 0.0), (60, 0.5),
                                      (10, 0.1)]):
        s.submit(req(k, steps, arr))
    got = s.admit(now_s=0.05, t=0)
    assert [(slot, r.rid) for slot, r in got] == [(0, 0), (1, 1)]
    assert s.admit(now_s=1.0, t=0) == []          # no free slot
    assert s.finished(t=19) == []
    done = s.finished(t=20)                        # request 1 ran 20 steps
    assert [(slot, r.rid, t0) for slot, r, t0, _ in done] == [(1, 1, 0)]
    got = s.admit(now_s=0.2, t=20)  # request 3 arrived at 0.1, 2 not yet
    assert [(slot, r.rid) for slot, r in got] == [(1, 3)]
    assert s.finished(t=29) == [] and len(s.finished(t=30)) == 1
    got = s.admit(now_s=0.6, t=40)
    assert [(slot, r.rid) for slot, r in got] == [(1, 2)]


def test_fifo_order_and_slot_reuse():
    s = Scheduler(3)
    for k in range(7):
        s.submit(req(k, 10 * (k + 1)))
    t, order, slots_seen = 0, [], set()
    while s.queue or s.busy:
        for slot, r in s.admit(now_s=0.0, t=t):
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
