"""Continuous batching on the GPU: requests admitted into slots between
windows, slots reused, every result identical to the same request run
alone, and one request against the oracle (same Philox stream)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
from flybrain_amd import connectome as C
from flybrain_amd.device import DeviceNet
from flybrain_amd.serve import Server, SimRequest

pytestmark = pytest.mark.gpu


def requests(n, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        out.append(SimRequest(
            rid=k, steps=int(rng.integers(25, 260)),
            drive=rng.choice(n, int(rng.integers(5, 40)), replace=False),
            rate_hz=float(rng.uniform(200, 1500)),
            silenced=rng.choice(n, int(rng.integers(0, 12)), replace=False),
            seed=int(rng.integers(0, 2**63))))
    return out


@pytest.fixture(scope="module")
def net():
    conn = C.synthetic(1500, None, mean_out=60, count_exp=1.15, n_groups=6,
                       seed=31)
    return DeviceNet(conn)


@pytest.mark.parametrize("road,slots,count,shard,small",
                         [("push", 6, 24, 64, ()),
                          ("pull", 6, 24, 64, None),
                          ("pull", 70, 24, 64, None),
                          ("pull", 128, 150, 64, None),
                          ("push", 40, 60, 16, ()),
                          ("push", 40, 60, 16, None),
                          ("push", 9, 40, 16, (1, 2, 2))])
def test_result_is_the_same_alone_and_in_the_mix(net, road, slots, count,
                                                 shard, small):
    """More requests than slots: slots (and whole shards, launched only
    while they hold a live request, each on its own stream) are freed and
    reused mid-run; small shards take the lowest slots."""
    reqs = requests(net.n, count, seed=slots)
    srv = Server(net, slots=slots, road=road, shard=shard, small=small)
    sizes = [sh.size for sh in srv.shards]
    assert sum(sizes) == slots and max(sizes) <= shard
    if small is None and road == "push":
        assert sizes[:8] == [1, 1, 1, 1, 4, 4, 4, 4]  # (then 8s, as fit)
    elif small:
        assert sizes[:len(small)] == list(small)
    else:
        assert len(sizes) == -(-slots // min(shard, slots))
    mix, stats = srv.serve([SimRequest(**vars(r)) for r in reqs],
                           realtime=False)
    assert stats["requests"] == len(reqs)
    got = {r.rid: r for r in mix}
    assert sorted(got) == [r.rid for r in reqs]
    assert len({r.slot for r in mix}) == min(slots, len(reqs))  # reused
    assert any(r.counts.sum() > 0 for r in mix)
    for r in reqs[:6] + reqs[-4:]:
        alone, _ = Server(net, slots=1, road="push").serve(
            [SimRequest(**vars(r))], realtime=False)
        np.testing.assert_array_equal(got[r.rid].counts, alone[0].counts,
                                      err_msg="request %d" % r.rid)


def test_served_request_matches_the_oracle(net):
    """A request admitted mid-run (its slot restarted at a later clock)
    gives the counts of the oracle running it from step 0."""
    reqs = requests(net.n, 5, seed=99)
    res, _ = Server(net, slots=2, road="pull").serve(
        [SimRequest(**vars(r)) for r in reqs], realtime=False)
    late = max(res, key=lambda r: r.admitted_s)
    r = reqs[late.rid]
    conn = net.conn
    ora = O.OracleEngine(conn, 1, seeds=[r.seed], trial_ids=[0],
                         no_refractory=np.unique(r.drive))
    if len(r.silenced):
        ora.set_masks(out=list(r.silenced))
    rates = np.zeros((1, conn.n), np.float32)
    rates[0, r.drive] = r.rate_hz
    counts = np.zeros(conn.n, np.int64)
    for _ in range(r.steps):
        counts += ora.step_rates(rates)[0].astype(np.int64)
    np.testing.assert_array_equal(late.counts, counts)
    assert counts.sum() > 0


@pytest.mark.parametrize("road,slots", [("push", 1), ("pull", 3)])
def test_graph_replay_serves_the_same(net, road, slots):
    """Whole graphs replayed while no slot is free give the results of
    window-by-window launches."""
    reqs = requests(net.n, 8, seed=5)
    eager, _ = Server(net, slots=slots, road=road).serve(
        [SimRequest(**vars(r)) for r in reqs], realtime=False)
    srv = Server(net, slots=slots, road=road, max_drive=64, graph_windows=2)
    graph, st = srv.serve([SimRequest(**vars(r)) for r in reqs],
                          realtime=False)
    assert srv._graph is not None, "the graph road was not taken"
    a = {r.rid: r.counts for r in eager}
    for r in graph:
        np.testing.assert_array_equal(r.counts, a[r.rid],
                                      err_msg="request %d" % r.rid)


def _bad_requests(n):
    base = dict(steps=80, drive=np.array([1, 2, 3]), rate_hz=300.0,
                silenced=np.zeros(0, np.int64), seed=5)
    return [SimRequest(rid=1000, **{**base, "drive": np.array([1, n])}),
            SimRequest(rid=1001, **{**base, "rate_hz": 20_000.0}),
            SimRequest(rid=1002, **{**base, "steps": 0}),
            SimRequest(rid=1003, **{**base, "drive": np.arange(300)}),
            SimRequest(rid=1004, **{**base, "silenced": np.array([-3])})]


@pytest.mark.parametrize("road,slots,graph_windows", [("push", 6, 0),
                                                      ("pull", 6, 0),
                                                      ("push", 5, 2)])
def test_bad_requests_cancel_and_eviction_leave_the_rest_unchanged(
        net, road, slots, graph_windows):
    """Rejected requests come back as errors without reaching a slot; a
    queued request is cancelled (nothing ran), a running one is cancelled
    with the counts of the steps it ran, another is evicted and later
    resumed into whatever slot is free.  Every other request's counts are
    bit-identical to a clean run of the same requests; the cancelled one
    equals the request run alone for the steps it ran, and the evicted one
    equals its uninterrupted result."""
    reqs = requests(net.n, 16, seed=77)
    for r in reqs:
        r.steps = max(r.steps, 120)
    clean, _ = Server(net, slots=slots, road=road,
                      graph_windows=graph_windows).serve(
        [SimRequest(**vars(r)) for r in reqs], realtime=False)
    want = {r.rid: r.counts for r in clean}

    srv = Server(net, slots=slots, road=road, graph_windows=graph_windows)
    seen = {"w": 0}
    plan = {}

    def on_window(s):
        w = seen["w"]
        seen["w"] += 1
        live = sorted((req.rid for req, _, _ in s.sched.active.values()))
        if w == 1:
            plan["queued"] = s.sched.queue[-1].rid
            s.cancel(plan["queued"])
        elif w == 3:
            plan["cancel"] = live[0]
            s.cancel(live[0])
        elif w == 4:
            plan["evict"] = live[-1]
            s.evict(live[-1])
        elif w == 9:
            assert plan["evict"] in s.suspended
            s.resume(plan["evict"])

    mix = [SimRequest(**vars(r)) for r in reqs] + _bad_requests(net.n)
    got, stats = srv.serve(mix, realtime=False, on_window=on_window)
    res = {r.rid: r for r in got}
    assert not srv.suspended
    for k in range(1000, 1005):
        assert res[k].error and res[k].counts is None, k
    assert stats["rejected"] == 5 and stats["cancelled"] == 2
    q = res[plan["queued"]]
    assert q.cancelled and q.steps_done == 0 and q.counts.sum() == 0
    c = res[plan["cancel"]]
    assert c.cancelled and 0 < c.steps_done < c.steps
    one = next(r for r in reqs if r.rid == plan["cancel"])
    alone, _ = Server(net, slots=1, road="push").serve(
        [SimRequest(**{**vars(one), "steps": c.steps_done})], realtime=False)
    np.testing.assert_array_equal(c.counts, alone[0].counts)
    for rid, counts in want.items():
        if rid in (plan["queued"], plan["cancel"]):
            continue
        assert res[rid].ok
        np.testing.assert_array_equal(res[rid].counts, counts,
                                      err_msg="request %d" % rid)


def test_shard_stream_context_sets_and_restores_the_current_stream(net):
    """A shard's stream context makes its stream torch's current one (the
    launches and torch's own work inside go there) and puts the previous
    one back, also when the block raises."""
    srv = Server(net, slots=3, road="push", shard=2, small=())
    before = torch.cuda.current_stream()
    for sh in srv.shards:
        with sh.on():
            assert torch.cuda.current_stream() == sh.stream
        assert torch.cuda.current_stream() == before
    with pytest.raises(KeyError):
        with srv.shards[0].on():
            raise KeyError("inside")
    assert torch.cuda.current_stream() == before
