"""Many independent trials in one state block, stepped in windows and
replayed from a captured HIP graph.

The reference batches trials that share one drive, one silencing and one
generator (lab/flybrain/sim.py:384-410, :1062).  Here every trial has its
own Philox seed, drive table entries and m_out masks; trials never
interact, so trial ``b`` of a batch produces the same raster as the same
trial run alone (tests check this bit for bit).

One window is the road's launches: on the push road update (W steps,
drawing the drive table's Philox kicks itself) and scatter; on the pull
road (64..1,024 trials, the default there) a drive-table launch and one
fused gather+update launch.
``graph_windows`` windows are captured into one graph.  The step
counter lives in device memory, so the same graph serves every replay.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from .device import DeviceNet, DeviceSim, DriveTable
from .workloads import Workload


class BatchRunner:
    def __init__(self, net: DeviceNet, wl: Workload, params=None,
                 raster_len=0, counts=True, pop_counts=False, window=20,
                 road="auto", fused_drive=True, pull_tile=None, gate=False):
        """``road="tuned"`` takes the road, pull-road workgroup shape and
        input gating that :func:`flybrain_amd.autotune.tune` measured
        fastest for this connectome, batch and workload (timed once, then
        cached)."""
        if not 1 <= window <= L.MAX_WINDOW:
            raise ValueError("window must be 1..%d" % L.MAX_WINDOW)
        if road == "tuned":
            from .autotune import tune
            best = tune(net, wl)["best"]
            road, pull_tile = best["road"], best["pull_tile"]
            gate = bool(best.get("gate", False))
        self.wl = wl
        self.window = int(window)
        sim = DeviceSim(net, wl.batch, params=params, seeds=wl.seeds,
                        trial_ids=np.zeros(wl.batch, np.uint32),
                        counts=counts, pop_counts=pop_counts,
                        raster_len=raster_len, road=road, pull_tile=pull_tile,
                        gate=gate)
        self.sim = sim
        # per-trial refractory exemption of driven neurons; the counter
        # starts at refrac_steps = 0 for them (sim.py:371-388)
        if wl.n_drive:
            sim.exempt_pairs(wl.drive_neuron, wl.drive_trial)
        if wl.out_neuron.size:
            sim.set_flag_pairs(L.OUT_OFF, wl.out_neuron, wl.out_trial)
        self.table = DriveTable(sim, wl.drive_neuron, wl.drive_trial,
                                wl.drive_group, wl.rate_hz, wl.seg_steps) \
            if wl.n_drive else None
        # the update launches draw the table's kicks themselves: one launch
        # less per window (the same bits as the drive-table launch)
        self.fused = bool(fused_drive and self.table is not None
                          and sim.road == "push"
                          and self.table.n_groups <= 32767)
        if self.fused:
            sim.fuse_drive(self.table)
        self._graph = None
        self._graph_windows = 0

    @property
    def t(self):
        return self.sim.t

    def _window(self, W):
        s = self.sim
        if self.table is not None and not self.fused:
            s.drive_table(self.table, W)
        s.step_launch(W)

    def capture(self, graph_windows=50):
        """Capture ``graph_windows`` full windows into one HIP graph.

        A warm-up window runs first (on a side stream) and the state is put
        back afterwards, as Engine.use_graphs does (sim.py:680-697).
        """
        s = self.sim
        names = ["v", "g", "st", "ring", "sbits", "ev_count", "clock"]
        names += [k for k in ("counts", "pop_counts", "raster", "rec", "act",
                              "due", "calm", "rec_ids")
                  if getattr(s, k) is not None]
        saved = {k: getattr(s, k).clone() for k in names}
        t0 = s.t
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self._window(self.window)
        torch.cuda.current_stream().wait_stream(side)
        g = torch.cuda.CUDAGraph()
        try:
            # one clock launch per replay, not one per window
            with torch.cuda.graph(g), s.deferred_clock():
                for _ in range(graph_windows):
                    self._window(self.window)
        finally:
            for k, val in saved.items():
                getattr(s, k).copy_(val)
            s.t = t0
        self._graph = g
        self._graph_windows = int(graph_windows)
        self._graph_version = s.version
        return g

    def advance(self, steps, use_graph=True):
        """Run ``steps`` steps: whole graphs where they fit, then windows."""
        left = int(steps)
        W = self.window
        if use_graph and self._graph is not None:
            if self._graph_version != self.sim.version:
                self.capture(self._graph_windows)
            per = W * self._graph_windows
            while left >= per:
                self._graph.replay()
                self.sim.t += per
                left -= per
        while left > 0:
            w = min(W, left)
            self._window(w)  # the road's launch advances the host clock
            left -= w

    def counts(self):
        """Spike counts per trial and neuron, int64 [batch, n]."""
        return self.sim.view(self.sim.counts).cpu().numpy().astype(np.int64)
