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

Around the loop, as the reference's job protocol has it (Submit / Cancel,
lab/flybrain/ipc.py:130-152; a failing job reported without stopping the
worker, lab/flybrain/worker.py:495-501; cancellation at a bin boundary with
the result truncated to what ran, lab/flybrain/jobs.py:348-412):

  * every request is checked when it is submitted (steps, drive and
    silenced indices, drive size, rate); a bad one comes back as a result
    with ``error`` set and never reaches a slot or the device;
  * ``cancel(rid)`` takes effect at the next window boundary the serving
    loop reaches (it launches up to ``batch_windows`` windows at a time
    while no request ends and no command waits): a queued
    request is dropped, a running one is read out with the counts of the
    steps it ran (``cancelled``, ``steps_done``) and its slot freed;
  * ``evict(rid)`` saves a running request's trial state (its column of
    v, g, refractory counter, flags, counts and the input in flight) and
    frees the slot; ``resume(rid)`` admits it again -- into any free slot,
    of any shard -- and it continues bit for bit as if never interrupted.
"""

from __future__ import annotations

import contextlib
import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L
from .device import NO_END, DeviceNet, DeviceSim, check_rates_hz, pick_road
from .device import _stream

# torch's current-stream slot read and set by stream id: what
# torch.cuda.stream() does for a stream of the current device, without the
# Stream and device objects it builds on every entry (~6 us; the serving
# loop enters a shard's stream once per shard and pass)
_get_cur = getattr(torch._C, "_cuda_getCurrentStream", None)
_set_cur = getattr(torch._C, "_cuda_setStream", None)


class _OnStream:
    """torch's current stream on ``stream``'s device set to ``stream`` for
    the block, the previous one restored after it."""

    __slots__ = ("sid", "dev", "typ", "prev")

    def __init__(self, stream):
        self.sid = stream.stream_id
        self.dev = stream.device_index
        self.typ = stream.device_type

    def __enter__(self):
        self.prev = _get_cur(self.dev)
        _set_cur(stream_id=self.sid, device_index=self.dev,
                 device_type=self.typ)
        return self

    def __exit__(self, *exc):
        p = self.prev
        _set_cur(stream_id=p[0], device_index=p[1], device_type=p[2])
        return False


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
    counts: np.ndarray  # int64 [n]: spikes per neuron over the steps run
    steps: int
    arrival_s: float
    admitted_s: float
    done_s: float
    slot: int
    # steps actually run: ``steps`` unless cancelled (then what ran before
    # the cancel took effect, the reference's truncation, jobs.py:348-364)
    steps_done: int = -1
    cancelled: bool = False
    error: str | None = None  # set for a rejected request (counts None)

    def __post_init__(self):
        if self.steps_done < 0:
            self.steps_done = self.steps if self.error is None else 0

    @property
    def ok(self):
        return self.error is None and not self.cancelled

    @property
    def latency_s(self):
        """Completion latency: result on the host minus arrival."""
        return self.done_s - self.arrival_s


def check_request(req, n, max_drive):
    """Why ``req`` cannot run on an ``n``-neuron network with ``max_drive``
    drive slots, or None.  Checked when the request is submitted, before
    any slot or device buffer is touched."""
    try:
        steps = int(req.steps)
        if steps != req.steps or steps < 1:
            return "steps must be a positive integer (got %r)" % (req.steps,)
        for name in ("drive", "silenced"):
            a = np.asarray(getattr(req, name))
            if a.size and (a.ndim != 1 or not np.issubdtype(a.dtype,
                                                            np.integer)):
                return "%s must be a 1-D array of neuron indices" % name
            if a.size and (int(a.min()) < 0 or int(a.max()) >= n):
                return "%s index outside [0, %d)" % (name, n)
        k = np.unique(np.asarray(req.drive, np.int64)).size
        if k > max_drive:
            return "drives %d neurons, more than max_drive = %d" % (k,
                                                                    max_drive)
        rate = float(req.rate_hz)
        if not np.isfinite(rate):
            return "rate_hz must be a finite number (got %r)" % (req.rate_hz,)
        check_rates_hz(np.array([rate]), "request %r" % (req.rid,))
        seed = int(req.seed)
        if not 0 <= seed < (1 << 64):
            return "seed outside [0, 2^64)"
    except (TypeError, ValueError) as e:  # RateTooHigh is a ValueError
        return str(e)
    return None


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
        # evicted requests asked back: admitted before new arrivals
        self.resumed = deque()  # (request, steps_done, admitted_s)

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
        """Resumed requests first, then arrived ones, into the lowest free
        slots.  Returns (slot, request, steps_done) with steps_done 0 for a
        fresh start; a resumed request keeps its own clock (t_start =
        t - steps_done)."""
        out = []
        while self.free and self.resumed:
            req, done, adm = self.resumed.popleft()
            slot = min(self.free)
            self.free.remove(slot)
            self.active[slot] = (req, int(t) - int(done), adm)
            out.append((slot, req, int(done)))
        while self.free and self.queue and self.queue[0].arrival_s <= now_s:
            req = self.queue.popleft()
            slot = min(self.free)
            self.free.remove(slot)
            self.active[slot] = (req, int(t), float(now_s))
            out.append((slot, req, 0))
        return out

    def release(self, slot):
        """Take a running request out of its slot (cancel, evict); returns
        (request, t_start, admitted_s)."""
        got = self.active.pop(slot)
        self.free.append(slot)
        return got

    def drop_queued(self, rid):
        """Remove a request that has not started; True if it was queued."""
        for q in (self.queue,):
            for k, req in enumerate(q):
                if req.rid == rid:
                    del q[k]
                    return True
        for k, (req, _, _) in enumerate(self.resumed):
            if req.rid == rid:
                del self.resumed[k]
                return True
        return False

    def slot_of(self, rid):
        for slot, (req, _, _) in self.active.items():
            if req.rid == rid:
                return slot
        return None

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

    @property
    def pending(self):
        return bool(self.queue or self.resumed)


def bytes_per_trial(n, road):
    """HBM held per trial slot: v, g, vhold (12 B), st, fl (2 B), int32
    ring (160 B), kicks and counts (8 B), events (6 B) and, on the pull
    road, 40 steps of spike records (5 B) -- per neuron."""
    per = 12 + 2 + 160 + 8 + 6 + (5 if road == "pull" else 0)
    return per * int(n)


class _Shard:
    """One DeviceSim of ``size`` slots and its drive.

    Slot s draws its kicks at rate ``_rate[s]`` (``_rate[size]`` = 0).  On
    the push road the update launch draws them itself (fused drive: a
    group index per neuron-trial, the slot's for its driven neurons, -1
    elsewhere; one launch per window).  On the pull road a drive table
    with a fixed layout feeds a drive launch: slot s owns entries
    [s * max_drive, (s + 1) * max_drive) and rate group s; group ``size``
    has rate 0 and pads the unused entries.  Admission writes one slot's
    part, retirement zeroes its rate, and no pointer or size ever changes
    (a captured graph stays valid).
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
        self.fused = self.sim.road == "push"
        if self.fused:
            self._groups = torch.full((self.sim.nb,), -1, dtype=torch.int16,
                                      device=dev)
            self._slot_idx = [None] * self.size  # driven neuron-trials
            self.sim.fuse_drive_groups(self._groups, self._rate,
                                       self.size + 1)
        self._clock = None  # the open deferred-clock run, if any
        # the shard's own stream (default priority), after its set-up
        self.stream = torch.cuda.Stream(device=dev)
        self.stream.wait_stream(torch.cuda.current_stream(dev))

    def on(self):
        """Run on the shard's stream: the work of different shards overlaps
        on the GPU, a shard's own stays in order."""
        if _get_cur is None or _set_cur is None:
            return torch.cuda.stream(self.stream)
        return _OnStream(self.stream)

    def set_slot_drive(self, slot, drive, rate_hz):
        if self.fused:
            old = self._slot_idx[slot]
            if old is not None:
                self._groups[old] = -1
            idx = torch.from_numpy(drive.astype(np.int64) * self.sim.bstride
                                   + slot).to(self.sim.device)
            self._groups[idx] = slot
            self._slot_idx[slot] = idx
        else:
            md = self.max_drive
            r0 = slot * md
            rows = self._ent_host[r0:r0 + md]
            rows[:, 0] = slot
            rows[:, 1] = self.size
            rows[:drive.size, 0] = (drive.astype(np.int64) * self.sim.bstride
                                    + slot)
            rows[:drive.size, 1] = slot
            self._entries[2 * r0:2 * (r0 + md)].copy_(
                torch.from_numpy(rows.view(np.int32).reshape(-1)))
        self._rate_host[slot] = np.float32(rate_hz)

    def push_rates(self):
        self._rate.copy_(torch.from_numpy(self._rate_host))

    def defer_clock(self):
        """Windows launched from now on leave the device clock alone (one
        tick when the run is closed, not one launch per window)."""
        if self._clock is None or self.sim.struct.clock_ahead >= (1 << 30):
            self.close_clock()
            self._clock = self.sim.deferred_clock()
            self._clock.__enter__()

    def close_clock(self):
        run, self._clock = self._clock, None
        if run is not None:
            run.__exit__(None, None, None)

    def window(self, W):
        s = self.sim
        if not self.fused:
            L.check(s.lib.fis_drive_table(
                s._sp, ctypes_ptr(self._entries), self.size * self.max_drive,
                ctypes_ptr(self._rate), self.size + 1, 1, 1 << 30, W,
                _stream()), "fis_drive_table")
        s.step_launch(W)


class Server:
    """Slots for SimRequests, held in shards of ``shard`` trials.

    Each shard is its own DeviceSim with its own HIP stream (default
    priority), so the windows of different shards overlap on the GPU; a
    window launches only the shards that hold a live request, so the GPU
    time of a window follows the live load rather than the capacity.  The
    lowest slots are ``small`` shards (by default, on the push road, four
    each of one, four and eight slots): the scheduler fills the lowest free
    slots first, so at low load requests run nearly as fast as alone (a
    20-step push window of one trial takes 22 us, of 16 trials 94 us), and
    the shards of ``shard`` slots take the load beyond them.  Measured on
    the full connectome (p50 / p99 ms at 200 and 20 requests/s): no small
    shards 36.0 / 74.9 and 34.9 / 86.2; four of one and four of four 15.5 /
    56.9 and 7.6 / 19.8; with four of eight as well 14.6 / 49.0 and 7.7 /
    20.0.  The shard of 16 trials (push road, one launch per window)
    measured best among uniform shards: p50 33 ms and p99 73 ms at 200
    requests/s, against 60 / 155 ms with 32 and 115 / 266 ms with 64 (pull
    road).

    ``cancel``, ``evict`` and ``resume`` may be called from any thread (or
    from the ``on_window`` / ``on_result`` callbacks of :meth:`serve`);
    they take effect at the next window boundary the serving loop reaches
    (at most ``batch_windows`` windows later).
    """

    def __init__(self, net: DeviceNet, slots=256, window=20, road="auto",
                 hbm_fraction=0.5, max_drive=None, graph_windows=0,
                 shard=16, batch_windows=4, small=None):
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
        # the small shards first (the scheduler fills the lowest free slots,
        # so a lightly loaded server runs its requests in them), then shards
        # of ``shard`` slots.  By default four each of one, four and eight
        # slots on the push road; none on the pull road, where every shard
        # pads to 64 trials
        if small is None:
            small = ((1,) * 4 + (4,) * 4 + (8,) * 4 if self.road == "push"
                     else ())
        sizes, left = [], self.slots
        for k in small:
            if 0 < int(k) <= left:
                sizes.append(int(k))
                left -= int(k)
        sizes += [self.shard] * (left // self.shard)
        if left % self.shard:
            sizes.append(left % self.shard)
        self.shards = [_Shard(net, k, pick_road(k, road), self.max_drive)
                       for k in sizes]
        # slot -> (shard, slot in the shard)
        self._shard_of = np.repeat(np.arange(len(sizes)), sizes)
        self._local_of = np.concatenate([np.arange(k) for k in sizes])
        self.sched = Scheduler(self.slots)
        self.t = 0  # the server's clock: steps run since it started
        self._graph = None
        self.graph_windows = int(graph_windows)
        # windows launched per loop pass when no request can end inside
        # them and no command waits (each pass schedules once)
        self.batch_windows = max(1, int(batch_windows))
        self._commands = deque()  # (what, rid), applied between windows
        self._lock = threading.Lock()
        # evicted requests: rid -> (request, trial state, admitted_s)
        self.suspended = {}
        self._results = None
        self._start = None
        self._on_result = None

    @staticmethod
    def capacity(net, road="auto", hbm_fraction=0.5):
        free, _ = torch.cuda.mem_get_info(net.device)
        per = bytes_per_trial(net.n, "pull" if road != "push" else "push")
        return int(free * hbm_fraction // per)

    def _where(self, slot):
        return (self.shards[int(self._shard_of[slot])],
                int(self._local_of[slot]))

    # ------------------------------------------------------------ control
    def cancel(self, rid):
        """Stop request ``rid`` at the next window boundary: dropped if it
        has not started, else its result holds the counts of the steps
        that ran (``cancelled``, ``steps_done``)."""
        with self._lock:
            self._commands.append(("cancel", rid))

    def evict(self, rid):
        """Take running request ``rid`` out of its slot at the next window
        boundary, keeping its trial state in ``suspended`` (no result)."""
        with self._lock:
            self._commands.append(("evict", rid))

    def resume(self, rid):
        """Admit suspended request ``rid`` again (before new arrivals, into
        any free slot); it continues from where it was evicted."""
        with self._lock:
            self._commands.append(("resume", rid))

    def _apply_commands(self, now):
        with self._lock:
            cmds = list(self._commands)
            self._commands.clear()
        for what, rid in cmds:
            if what == "resume":
                if rid in self.suspended:
                    req, state, adm = self.suspended[rid]
                    self.sched.resumed.append((req, state["steps_done"],
                                               adm))
                continue
            slot = self.sched.slot_of(rid)
            if what == "cancel":
                if slot is not None:
                    req, t0, adm = self.sched.release(slot)
                    sh, local = self._where(slot)
                    with sh.on():
                        counts = sh.sim.trial_counts(local)
                    self._emit(SimResult(
                        req.rid, counts, req.steps, req.arrival_s, adm,
                        time.perf_counter() - self._start, slot,
                        steps_done=self.t - t0, cancelled=True))
                    self._retire(slot)
                elif rid in self.suspended:  # counts up to the eviction
                    req, state, adm = self.suspended.pop(rid)
                    for k, (r, _, _) in enumerate(self.sched.resumed):
                        if r.rid == rid:
                            del self.sched.resumed[k]
                            break
                    self._emit(SimResult(
                        rid, state["counts"].to(torch.int64).cpu().numpy(),
                        req.steps, req.arrival_s, adm,
                        time.perf_counter() - self._start, -1,
                        steps_done=state["steps_done"], cancelled=True))
                else:
                    req = next((r for r in self.sched.queue if r.rid == rid),
                               None)
                    if req is not None and self.sched.drop_queued(rid):
                        self._emit(SimResult(
                            rid, np.zeros(self.net.n, np.int64), req.steps,
                            req.arrival_s, now,
                            time.perf_counter() - self._start, -1,
                            steps_done=0, cancelled=True))
            elif what == "evict" and slot is not None:
                req, t0, adm = self.sched.release(slot)
                sh, local = self._where(slot)
                with sh.on():
                    state = sh.sim.trial_state(local)
                sh.stream.synchronize()  # any shard's stream may load it
                self.suspended[req.rid] = (req, state, adm)
                self._retire(slot)

    def _land(self, landing, start, wait=False):
        """Emit the results whose counts have reached the host (all of
        them with ``wait``, after a synchronisation); the rest stay."""
        left = []
        for (host, ev), req, adm, slot in landing:
            if wait or ev.query():
                self._emit(SimResult(req.rid, host.numpy(), req.steps,
                                     req.arrival_s, adm,
                                     time.perf_counter() - start, slot))
            else:
                left.append(((host, ev), req, adm, slot))
        return left

    def _emit(self, res):
        self._results.append(res)
        if self._on_result is not None:
            self._on_result(res)

    def _capture(self):
        """One HIP graph of ``graph_windows`` windows of every shard.  The
        drive tables, the per-trial clocks and the step counters are read
        from device memory, so the graph stays valid while requests come
        and go."""
        self._close_clocks()
        self._join_streams()
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

    def _close_clocks(self):
        for sh in self.shards:
            with sh.on():
                sh.close_clock()

    def _join_streams(self):
        """The current stream waits for every shard's work."""
        cur = torch.cuda.current_stream()
        for sh in self.shards:
            cur.wait_stream(sh.stream)

    def _fork_streams(self):
        """Every shard's stream waits for the current stream's work."""
        cur = torch.cuda.current_stream()
        for sh in self.shards:
            sh.stream.wait_stream(cur)

    def _advance(self):
        """Run windows of every shard holding a live request: one, or up to
        ``batch_windows`` when no request ends inside them and no control
        command is waiting -- or, when in addition no slot is free (nothing
        could be admitted meanwhile) and no shard still consumes a resumed
        trial's input from its ring, one whole captured graph of windows.
        Eager windows run in each shard's open deferred-clock run."""
        per = self.graph_windows * self.window
        left = min(t0 + req.steps for req, t0, _ in
                   self.sched.active.values()) - self.t
        if (per and not self.sched.free and not self._commands
                and all(sh.sim.t >= sh.sim.ring_until for sh in self.shards)
                and left >= per):
            if self._graph is None:
                self._capture()
            self._close_clocks()
            self._join_streams()
            self._graph.replay()
            self._fork_streams()
            for sh in self.shards:
                sh.sim.t += per
            self.t += per
            return self.graph_windows
        k = 1
        if not self._commands:
            k = max(1, min(self.batch_windows, left // self.window))
        for sh in self.shards:  # (a shard's own clock moves with its windows)
            if sh.live:
                with sh.on():
                    sh.defer_clock()
                    for _ in range(k):
                        sh.window(self.window)
        self.t += k * self.window
        return k

    # ------------------------------------------------------------ slots
    def _admit(self, admits):
        """Start (or resume) admitted requests in their slots.  Requests
        were checked at submission, so nothing here can fail on a bad
        request."""
        by_shard = {}
        for slot, req, done in admits:
            drv = np.unique(np.asarray(req.drive, np.int64))
            sh, local = self._where(slot)
            by_shard.setdefault(id(sh), (sh, []))[1].append(
                (local, req, drv, done))
        for sh, items in by_shard.values():
            with sh.on():
                self._admit_shard(sh, items)

    def _admit_shard(self, sh, items):
        s = sh.sim
        fresh = [(lo, r, d) for lo, r, d, done in items if done == 0]
        for local, req, drv, done in items:
            sh.set_slot_drive(local, drv, req.rate_hz)
            if done:  # a resumed request: its saved column goes back
                _, state, _ = self.suspended.pop(req.rid)
                s.load_trial_state(local, state)
                # (the saved tensors may belong to another shard's stream:
                # they are not freed before this stream has read them)
                sh.stream.synchronize()
        if fresh:
            out_ne, out_tr, free_ne, free_tr = [], [], [], []
            for local, req, drv in fresh:
                sil = np.unique(np.asarray(req.silenced, np.int64))
                out_ne.append(sil)
                out_tr.append(np.full(sil.size, local))
                free_ne.append(drv)
                free_tr.append(np.full(drv.size, local))
            s.restart_trials([lo for lo, _, _ in fresh],
                             [int(r.seed) for _, r, _ in fresh],
                             [s.t + int(r.steps) for _, r, _ in fresh],
                             out_neurons=np.concatenate(out_ne),
                             out_trials=np.concatenate(out_tr),
                             free_neurons=np.concatenate(free_ne),
                             free_trials=np.concatenate(free_tr))
        sh.push_rates()
        sh.live += len(items)

    def _retire(self, slot):
        sh, local = self._where(slot)
        with sh.on():
            sh._rate_host[local] = 0.0
            sh.sim.set_trial_end([local], [0])
            sh.live -= 1
            sh.push_rates()

    # ------------------------------------------------------------ loop
    def serve(self, requests, realtime=True, on_result=None, on_window=None):
        """Serve ``requests`` (arrival_s relative to the call) and return
        (results in completion order, stats).  With ``realtime`` the loop
        admits a request only once its arrival time has passed on the wall
        clock; otherwise arrival times are ignored (all queued at 0).

        A request that fails :func:`check_request` is answered at once
        with an ``error`` result.  ``on_window(server)`` is called at every
        window boundary, before commands are applied and requests
        admitted.  Requests evicted and not resumed stay in
        ``suspended`` when the call returns."""
        self._results = results = []
        self._on_result = on_result
        self._start = start = time.perf_counter()
        reqs = sorted(requests, key=lambda r: r.arrival_s)
        for r in reqs:
            why = check_request(r, self.net.n, self.max_drive)
            if why is not None:
                self._emit(SimResult(r.rid, None, r.steps, r.arrival_s,
                                     r.arrival_s, 0.0, -1, error=why))
                continue
            if not realtime:
                r.arrival_s = 0.0
            self.sched.submit(r)
        busy_windows = 0
        landing = []  # (counts copy in flight, request, admitted, slot)
        while (self.sched.pending or self.sched.busy or self._commands):
            if landing:  # (also while the loop waits for arrivals)
                landing = self._land(landing, start)
            now = time.perf_counter() - start
            if on_window is not None:
                on_window(self)
            self._apply_commands(now)
            admits = self.sched.admit(now, self.t)
            if admits:
                self._admit(admits)
            if not self.sched.busy:
                nxt = self.sched.next_arrival()
                if nxt is not None:
                    time.sleep(max(0.0, min(nxt - now, 0.001)))
                elif not self._commands:
                    break  # only suspended requests are left
                continue
            busy_windows += self._advance()
            # a finished request's counts are copied to the host on its
            # shard's stream, and its slot freed at once (whatever runs
            # there next is ordered after the copy); the result is emitted
            # once the copy has landed, so the loop never waits for it
            for slot, req, t0, adm in self.sched.finished(self.t):
                sh, local = self._where(slot)
                with sh.on():
                    landing.append((sh.sim.trial_counts_async(local), req,
                                    adm, slot))
                self._retire(slot)
            landing = self._land(landing, start)
        self._close_clocks()
        torch.cuda.synchronize()
        self._land(landing, start, wait=True)
        wall = time.perf_counter() - start
        self._on_result = None
        return results, summarize(results, wall, self.slots, busy_windows,
                                  self.window)


def summarize(results, wall, slots, windows, window):
    """Throughput and completion latency of the requests that ran to the
    end; rejected and cancelled ones are counted apart (cancelled steps
    still count as simulated work)."""
    done = [r for r in results if getattr(r, "ok", True)]
    lat = np.array([r.latency_s for r in done]) if done else np.zeros(1)
    steps = sum(r.steps_done if hasattr(r, "steps_done") else r.steps
                for r in results if getattr(r, "error", None) is None)
    return {"requests": len(done), "slots": int(slots),
            "rejected": sum(1 for r in results
                            if getattr(r, "error", None) is not None),
            "cancelled": sum(1 for r in results
                             if getattr(r, "cancelled", False)),
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


__all__ = ["SimRequest", "SimResult", "Scheduler", "Server", "NO_END",
           "bytes_per_trial", "check_request", "poisson_requests",
           "summarize"]
