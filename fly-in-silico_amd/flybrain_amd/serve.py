"""Serving: continuous batching of simulation requests on one GPU.

The reference serves one simulation job at a time: the worker takes the
next job, builds an engine, runs it to completion and writes the result
(lab/flybrain/worker.py:298-313, :454-505; cancellation and truncation in
lab/flybrain/jobs.py:348-412).  Here the requests share one resident batch
of trial slots:

  * capacity is chosen from free HBM: ``bytes_per_trial`` of state, delay
    ring and spike records per slot (``Server.capacity``);
  * requests that have arrived are admitted into free slots between two
    windows; the slot is reset on the device (``fis_reset_trials``), its
    own clock starts at 0 (Philox stream and rate schedule are the
    request's own, independent of when and where it runs) and its spikes
    are counted until its last step (``t_end``);
  * all admitted requests advance together, one window of 20 steps per
    launch, and a request whose last step has run is read out (its column
    of spike counts) and its slot freed for the next arrival;
  * slots are held in shards (64 trials each by default, one DeviceSim per
    shard) and a window launches only the shards holding a live request,
    so a partly filled server costs what its live shards cost.

Trials never interact, so a request's result is bit-for-bit the result of
the same request run alone; ``tests/test_gpu_serve.py`` checks this.
"""

from __future__ import annotations

import contextlib
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L
from .device import NO_END, DeviceNet, DeviceSim, check_rates_hz, pick_road


@dataclass
class SimRequest:
    """One simulation: ``steps`` steps of 0.1 ms with Poisson drive into
    ``drive`` at ``rate_hz``, m_out = 0 on ``silenced``, Philox ``seed``."""

    rid: int
    steps: int
    drive: np.ndarray
    rate_hz: float
    silenced: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    seed: int = 0
    arrival_s: float = 0.0


@dataclass
class SimResult:
    rid: int
    counts: np.ndarray  # int64 [n]: spikes per neuron over the request
    steps: int
    arrival_s: float
    admitted_s: float
    done_s: float
    slot: int

    @property
    def latency_s(self):
        """Completion latency: result on the host minus arrival."""
        return self.done_s - self.arrival_s


class Scheduler:
    """Queue and slot bookkeeping (no device work).

    ``submit`` queues requests in arrival order; ``admit`` moves arrived
    requests into free slots (lowest slot first) at global step ``t``;
    ``finished`` frees the slots whose request's last step is < ``t``.
    """

    def __init__(self, slots):
        if slots < 1:
            raise ValueError("need at least one slot")
        self.slots = int(slots)
        self.free = list(range(self.slots))
        self.active = {}  # slot -> (request, t_start, admitted_s)
        self.queue = deque()

    def submit(self, req):
        """Queue a request; the queue stays in arrival order (ties in
        submission order)."""
        if req.steps < 1:
            raise ValueError("request %r: steps must be >= 1" % (req.rid,))
        k = len(self.queue)
        while k > 0 and self.queue[k - 1].arrival_s > req.arrival_s:
            k -= 1
        self.queue.insert(k, req)

    def admit(self, now_s, t):
        out = []
        while self.free and self.queue and self.queue[0].arrival_s <= now_s:
            req = self.queue.popleft()
            slot = min(self.free)
            self.free.remove(slot)
            self.active[slot] = (req, int(t), float(now_s))
            out.append((slot, req))
        return out

    def finished(self, t):
        done = [s for s, (req, t0, _) in self.active.items()
                if t0 + req.steps <= t]
        out = []
        for s in sorted(done):
            req, t0, adm = self.active.pop(s)
            self.free.append(s)
            out.append((s, req, t0, adm))
        return out

    def next_arrival(self):
        return self.queue[0].arrival_s if self.queue else None

    @property
    def busy(self):
        return bool(self.active)


def bytes_per_trial(n, road):
    """HBM held per trial slot: v, g, vhold (12 B), st, fl (2 B), int32
    ring (160 B), kicks and counts (8 B), events (6 B) and, on the pull
    road, 40 steps of spike records (5 B) -- per neuron."""
    per = 12 + 2 + 160 + 8 + 6 + (5 if road == "pull" else 0)
    return per * int(n)


class _Shard:
    """One DeviceSim of ``size`` slots and its drive table.

    The drive table has a fixed layout: slot s owns entries
    [s * max_drive, (s + 1) * max_drive) and rate group s; group ``size``
    has rate 0 and pads the unused entries.  Admission writes one slot's
    range, retirement zeroes its rate, and the table's pointers and size
    never change (a captured graph stays valid).
    """

    def __init__(self, net, size, road, max_drive):
        self.size = int(size)
        self.sim = DeviceSim(net, self.size, counts=True, road=road,
                             seeds=np.zeros(self.size, np.uint64),
                             trial_ids=np.zeros(self.size, np.uint32))
        # idle slots count nothing; on the pull road, 64-slot words with no
        # live request are skipped
        self.sim.track_idle_words()
        self.sim.set_trial_end(np.arange(self.size), np.zeros(self.size))
        self.max_drive = md = int(max_drive)
        ent = np.zeros((self.size * md, 2), np.uint32)
        ent[:, 0] = np.arange(self.size * md) // md  # neuron 0, own trial
        ent[:, 1] = self.size
        self._ent_host = ent
        self._rate_host = np.zeros(self.size + 1, np.float32)
        dev = self.sim.device
        self._entries = torch.from_numpy(ent.view(np.int32).reshape(-1)).to(dev)
        self._rate = torch.zeros(self.size + 1, dtype=torch.float32,
                                 device=dev)
        self.live = 0

    def set_slot_drive(self, slot, drive, rate_hz):
        md = self.max_drive
        r0 = slot * md
        rows = self._ent_host[r0:r0 + md]
        rows[:, 0] = slot
        rows[:, 1] = self.size
        rows[:drive.size, 0] = drive.astype(np.int64) * self.sim.bstride + slot
        rows[:drive.size, 1] = slot
        self._entries[2 * r0:2 * (r0 + md)].copy_(
            torch.from_numpy(rows.view(np.int32).reshape(-1)))
        self._rate_host[slot] = np.float32(rate_hz)

    def push_rates(self):
        self._rate.copy_(torch.from_numpy(self._rate_host))

    def window(self, W):
        s = self.sim
        L.check(s.lib.fis_drive_table(
            s._sp, ctypes_ptr(self._entries), self.size * self.max_drive,
            ctypes_ptr(self._rate), self.size + 1, 1, 1 << 30, W,
            ctypes_stream()), "fis_drive_table")
        s.step_launch(W)


class Server:
    """Slots for SimRequests, held in shards of ``shard`` trials.

    Each shard is its own DeviceSim; a window launches only the shards
    that hold a live request, so the GPU time of a window follows the live
    load rather than the capacity.  The default shard of 16 trials (push
    road, one launch per window) measured best on the full connectome:
    p50 33 ms and p99 73 ms at 200 requests/s, against 60 / 155 ms with 32
    and 115 / 266 ms with 64 (pull road).
    """

    def __init__(self, net: DeviceNet, slots=256, window=20, road="auto",
                 hbm_fraction=0.5, max_drive=None, graph_windows=0,
                 shard=16):
        self.net = net
        cap = self.capacity(net, road, hbm_fraction)
        self.slots = int(min(slots, cap))
        if self.slots < 1:
            raise MemoryError("no HBM for a single trial slot")
        self.window = int(window)
        if not 1 <= self.window <= L.MAX_WINDOW:
            raise ValueError("window must be 1..%d" % L.MAX_WINDOW)
        self.shard = int(min(shard, self.slots))
        if self.shard < 1:
            raise ValueError("shard must hold at least one slot")
        self.road = pick_road(self.shard, road)
        self.max_drive = int(max_drive or 256)
        sizes = [self.shard] * (self.slots // self.shard)
        if self.slots % self.shard:
            sizes.append(self.slots % self.shard)
        self.shards = [_Shard(net, k, pick_road(k, road), self.max_drive)
                       for k in sizes]
        self.sched = Scheduler(self.slots)
        self.t = 0  # the server's clock: steps run since it started
        self._graph = None
        self.graph_windows = int(graph_windows)

    @staticmethod
    def capacity(net, road="auto", hbm_fraction=0.5):
        free, _ = torch.cuda.mem_get_info(net.device)
        per = bytes_per_trial(net.n, "pull" if road != "push" else "push")
        return int(free * hbm_fraction // per)

    def _where(self, slot):
        return self.shards[slot // self.shard], slot % self.shard

    def _capture(self):
        """One HIP graph of ``graph_windows`` windows of every shard.  The
        drive tables, the per-trial clocks and the step counters are read
        from device memory, so the graph stays valid while requests come
        and go."""
        g = torch.cuda.CUDAGraph()
        t0 = [sh.sim.t for sh in self.shards]  # capture only records
        try:
            with torch.cuda.graph(g), contextlib.ExitStack() as clocks:
                # one clock launch per shard and replay, not per window
                for sh in self.shards:
                    clocks.enter_context(sh.sim.deferred_clock())
                for _ in range(self.graph_windows):
                    for sh in self.shards:
                        sh.window(self.window)
        finally:
            for sh, t in zip(self.shards, t0):
                sh.sim.t = t
        self._graph = g

    def _advance(self):
        """Run one window of every shard holding a live request -- or, when
        no slot is free (nothing could be admitted meanwhile) and no request
        ends inside it, one whole captured graph of windows."""
        per = self.graph_windows * self.window
        if per and not self.sched.free:
            left = min(t0 + req.steps for req, t0, _ in
                       self.sched.active.values()) - self.t
            if left >= per:
                if self._graph is None:
                    self._capture()
                self._graph.replay()
                for sh in self.shards:
                    sh.sim.t += per
                self.t += per
                return self.graph_windows
        for sh in self.shards:
            if sh.live:
                sh.window(self.window)
        self.t += self.window
        return 1

    # ------------------------------------------------------------ slots
    def _admit(self, admits):
        by_shard = {}
        for slot, req in admits:
            drv = np.unique(np.asarray(req.drive, np.int64))
            check_rates_hz(np.array([req.rate_hz]), "request %r" % (req.rid,))
            if drv.size > self.max_drive:
                raise ValueError("request %r drives %d neurons > max_drive"
                                 % (req.rid, drv.size))
            sh, local = self._where(slot)
            by_shard.setdefault(id(sh), (sh, []))[1].append((local, req, drv))
        for sh, items in by_shard.values():
            s = sh.sim
            out_ne, out_tr, free_ne, free_tr = [], [], [], []
            for local, req, drv in items:
                sil = np.unique(np.asarray(req.silenced, np.int64))
                out_ne.append(sil)
                out_tr.append(np.full(sil.size, local))
                free_ne.append(drv)
                free_tr.append(np.full(drv.size, local))
                sh.set_slot_drive(local, drv, req.rate_hz)
            s.restart_trials([lo for lo, _, _ in items],
                             [int(r.seed) for _, r, _ in items],
                             [s.t + int(r.steps) for _, r, _ in items],
                             out_neurons=np.concatenate(out_ne),
                             out_trials=np.concatenate(out_tr),
                             free_neurons=np.concatenate(free_ne),
                             free_trials=np.concatenate(free_tr))
            sh.push_rates()
            sh.live += len(items)

    def _retire(self, slot):
        sh, local = self._where(slot)
        sh._rate_host[local] = 0.0
        sh.sim.set_trial_end([local], [0])
        sh.live -= 1
        sh.push_rates()

    # ------------------------------------------------------------ loop
    def serve(self, requests, realtime=True, on_result=None):
        """Serve ``requests`` (arrival_s relative to the call) and return
        (results in completion order, stats).  With ``realtime`` the loop
        admits a request only once its arrival time has passed on the wall
        clock; otherwise arrival times are ignored (all queued at 0)."""
        reqs = sorted(requests, key=lambda r: r.arrival_s)
        for r in reqs:
            if not realtime:
                r.arrival_s = 0.0
            self.sched.submit(r)
        results = []
        start = time.perf_counter()
        busy_windows = 0
        while self.sched.queue or self.sched.busy:
            now = time.perf_counter() - start
            admits = self.sched.admit(now, self.t)
            if admits:
                self._admit(admits)
            if not self.sched.busy:
                nxt = self.sched.next_arrival()
                if nxt is not None:
                    time.sleep(max(0.0, min(nxt - now, 0.001)))
                continue
            busy_windows += self._advance()
            for slot, req, t0, adm in self.sched.finished(self.t):
                sh, local = self._where(slot)
                counts = sh.sim.trial_counts(local)  # synchronises
                res = SimResult(req.rid, counts, req.steps, req.arrival_s,
                                adm, time.perf_counter() - start, slot)
                results.append(res)
                self._retire(slot)
                if on_result is not None:
                    on_result(res)
        torch.cuda.synchronize()
        wall = time.perf_counter() - start
        return results, summarize(results, wall, self.slots, busy_windows,
                                  self.window)


def summarize(results, wall, slots, windows, window):
    lat = np.array([r.latency_s for r in results]) if results else np.zeros(1)
    steps = sum(r.steps for r in results)
    return {"requests": len(results), "slots": int(slots),
            "wall_s": round(wall, 4),
            "trial_s_per_s": round(steps * 1e-4 / wall, 3) if wall else 0.0,
            "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
            "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
            "mean_ms": round(1e3 * float(lat.mean()), 3),
            "windows": int(windows), "window_steps": int(window)}


def poisson_requests(n, count, rate_req_s, seed=0, steps=(500, 20_000),
                     drive=(20, 200), drive_hz=(50.0, 250.0),
                     silenced=(0, 50)):
    """Requests of the serving workload: Poisson arrivals at
    ``rate_req_s``, duration U(500, 20,000) steps, 20-200 driven neurons
    at U(50, 250) Hz, 0-50 silenced (m_out = 0) neurons, own seed."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_req_s, size=count))
    out = []
    for k in range(count):
        nd = int(rng.integers(drive[0], drive[1] + 1))
        ns = int(rng.integers(silenced[0], silenced[1] + 1))
        out.append(SimRequest(
            rid=k, steps=int(rng.integers(steps[0], steps[1] + 1)),
            drive=rng.choice(n, nd, replace=False),
            rate_hz=float(rng.uniform(*drive_hz)),
            silenced=rng.choice(n, ns, replace=False),
            seed=int(rng.integers(0, 2**63)), arrival_s=float(t[k])))
    return out


def ctypes_ptr(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr())


def ctypes_stream():
    import ctypes
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


__all__ = ["SimRequest", "SimResult", "Scheduler", "Server", "NO_END",
           "bytes_per_trial", "poisson_requests", "summarize"]
