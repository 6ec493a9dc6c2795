"""Device state of one batch of LIF trials and the launches that advance it.

``DeviceNet`` holds the connectome (by presynaptic neuron) in HBM;
``DeviceSim`` holds the state of ``batch`` trials in the trial-minor layout
the kernels use (idx = neuron * bstride + trial) and enqueues the C-ABI
entry points of libfis.so on the current torch stream.  torch is used for
device memory and streams only; all arithmetic of the step is in the HIP
kernels.

Memory per neuron-trial: v, g (4 B each), st, fl (1 B each), sbits (4 B),
ring 40 x 4 B -- about 174 B; at 138,639 neurons x 256 trials that is
6.2 GB plus events, far inside one MI355X's 288 GB.
"""

from __future__ import annotations

import contextlib
import ctypes

import numpy as np
import torch

from . import _lib as L
from .params import derived


ROADS = ("push", "pull")
NO_END = 1 << 62  # t_end of a trial that never stops counting
PULL_MIN_BATCH = 64


def _nonempty(a):
    """``a``, or one zero of its dtype when it is empty."""
    return a if a.size else np.zeros(1, a.dtype)


def pick_road(batch: int, road: str = "auto") -> str:
    """``auto``: the pull road from 64 trials up (to 1,024), else push."""
    if road == "auto":
        return "pull" if PULL_MIN_BATCH <= batch <= L.PULL_MAX else "push"
    if road not in ROADS:
        raise ValueError("road %r not in %s" % (road, ROADS))
    if road == "pull" and batch > L.PULL_MAX:
        raise ValueError("the pull road holds at most %d trials"
                         % L.PULL_MAX)
    return road


def stride_for(batch: int, road: str = "push") -> int:
    """Trial stride.  Push road: next power of two below 64, next multiple
    of 64 above.  Pull road: next power of two, at least 64."""
    if batch < 1:
        raise ValueError("batch must be >= 1")
    if road == "pull":
        p = 64
        while p < batch:
            p *= 2
        return p
    if batch >= 64:
        return (batch + 63) // 64 * 64
    p = 1
    while p < batch:
        p *= 2
    return p


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


# torch's current HIP stream of the current device as a raw handle: the
# private getters cost ~0.1 us a call each, torch.cuda.current_stream()
# ~2.8 us (most of a one-step Engine.step's host time) and
# torch.cuda.current_device() ~0.8.  Every launch comes after the
# simulation's tensors were allocated on the device, so torch's GPU state
# is initialised by then.
_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)
_get_device = getattr(torch._C, "_cuda_getDevice", None)


def _stream():
    if _raw_stream is not None and _get_device is not None:
        return ctypes.c_void_p(_raw_stream(_get_device()))
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class DeviceNet:
    """The synapse matrix by presynaptic neuron, resident in HBM."""

    def __init__(self, conn, device="cuda"):
        self.conn = conn
        self.n = conn.n
        dev = torch.device(device)
        if dev.type == "cuda" and dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.offsets = torch.from_numpy(conn.offsets).to(self.device)
        # edge arrays of at least one element: a network without edges
        # still passes the library valid (never read) pointers
        self.cols = torch.from_numpy(_nonempty(conn.cols)).to(self.device)
        self.vals = torch.from_numpy(_nonempty(conn.vals)).to(self.device)
        self._layouts = {}

    # pull road: order neurons by neuropil group (see pull_layout).  Off by
    # default: measured on MI355X at 138,639 neurons, the identity order
    # (no dependent load of each tile's neuron numbers) was faster on every
    # workload (B = 256 sensory 0.902 -> 0.882 ms per window, B = 64
    # 0.179 -> 0.176, plain B = 256 gated 0.312 -> 0.281)
    locality = False

    def pull_layout(self, bstride):
        """The pull road's neuron order and its in-edge CSR by position.

        Units of max(1, 256 / bstride) aligned neurons (what one 256-lane
        group of the kernel covers, so that every line a workgroup writes
        stays inside one tile) are sorted by the neuropil group of their
        first neuron (stable; a partial last unit stays last).  Neurons of
        one group share most presynaptic partners (80 % of edges stay
        inside a group), so a run of consecutive tiles -- what one XCD
        takes -- reads the spike records of few presynaptic neurons.
        Without groups the order is the identity.  Built once per unit
        size; int32 offsets.
        """
        q = max(1, 256 // int(bstride))
        key = (q, bool(self.locality))
        got = self._layouts.get(key)
        if got is not None:
            return got
        c = self.conn
        if c.n_edges >= (1 << 31):
            raise ValueError("the pull road indexes edges with int32")
        n = c.n
        if c.groups is not None and self.locality:
            full = n // q
            unit_group = np.asarray(c.groups, np.int64)[:full * q:q]
            units = np.argsort(unit_group, kind="stable")
            order = (units[:, None] * q + np.arange(q)[None, :]).reshape(-1)
            order = np.concatenate([order, np.arange(full * q, n)])
        else:
            order = np.arange(n)
        # in-edges of neuron order[k] at positions in_off[k] .. in_off[k+1]
        pre = np.repeat(np.arange(n, dtype=np.int32), np.diff(c.offsets))
        by_post = np.argsort(c.cols, kind="stable")
        indeg = np.bincount(c.cols, minlength=n)
        nat = np.zeros(n + 1, np.int64)
        np.cumsum(indeg, out=nat[1:])
        lens = indeg[order]
        off = np.zeros(n + 1, np.int64)
        np.cumsum(lens, out=off[1:])
        take = by_post[np.repeat(nat[order] - off[:-1], lens)
                       + np.arange(c.n_edges)]
        dev = self.device
        got = (torch.from_numpy(order.astype(np.int32)).to(dev),
               torch.from_numpy(off.astype(np.int32)).to(dev),
               torch.from_numpy(_nonempty(pre[take])).to(dev),
               torch.from_numpy(_nonempty(c.vals[take])).to(dev))
        self._layouts[key] = got
        return got


class DeviceSim:
    """State of ``batch`` independent trials on one GPU."""

    def __init__(self, net: DeviceNet, batch: int, params=None, seeds=None,
                 trial_ids=None, no_refractory=None, spikes_out=False,
                 counts=False, pop_counts=False, raster_len=0, road="push",
                 pull_tile=None, gate=False):
        """``gate`` (pull road): skip the in-edge walk of tiles that no
        input can reach in a window (fis_sim.due); fixed for the
        simulation's life, chosen by measurement (flybrain_amd.autotune)."""
        self.lib = L.lib()
        self.road = pick_road(int(batch), road)
        # pull road: neuron-trials per workgroup (fixed for the simulation:
        # the act words are laid out by it)
        self.pull_tile = int(pull_tile or L.PULL_TILE)
        if self.pull_tile not in L.PULL_TILES:
            raise ValueError("pull_tile must be one of %s" % (L.PULL_TILES,))
        if net.device.type != "cuda":
            raise L.KernelUnavailable("the engine runs on a GPU only; got "
                                      "device %r" % (net.device,))
        self.net = net
        self.c = derived(params)
        self.n = n = net.n
        self.batch = B = int(batch)
        self.bstride = bs = stride_for(B, self.road)
        self.nb = nb = n * bs
        self.n_waves = int(self.lib.fis_n_waves(n, bs))
        self.n_blocks = int(self.lib.fis_n_blocks(n, bs))
        # a window's events plus one sink slot per wave segment, rounded up
        # so every segment starts on its own 128-byte line
        cap = L.MAX_WINDOW * (1 if bs >= 64 else 64 // bs) + 1
        self.ev_cap = (cap + L.EVENT_ALIGN - 1) // L.EVENT_ALIGN * L.EVENT_ALIGN
        dev = net.device
        self.device = dev
        z = dict(device=dev)
        self.v = torch.full((nb,), float(self.c["v0"]), dtype=torch.float32, **z)
        self.g = torch.zeros(nb, dtype=torch.float32, **z)
        # byte arrays padded to whole 32-bit words (fis_flag_pairs)
        nb4 = (nb + 3) // 4 * 4
        self.st = torch.zeros(nb4, dtype=torch.uint8, **z)[:nb]
        self.fl = torch.zeros(nb4, dtype=torch.uint8, **z)[:nb]
        # allocated up front: a captured graph holds this pointer
        self.vhold = torch.zeros(nb, dtype=torch.float32, **z)
        self.ring = torch.zeros(L.RING_LEN * nb, dtype=torch.int32, **z)
        self.sbits = torch.zeros(nb, dtype=torch.int32, **z)
        self.events = torch.zeros(self.n_waves * self.ev_cap * 2,
                                  dtype=torch.int64, **z)
        self.ev_count = torch.zeros(self.n_blocks * L.LINE_INTS,
                                    dtype=torch.int32, **z)
        self.clock = torch.zeros(1, dtype=torch.int64, **z)
        sd = np.zeros(bs, np.uint64)
        sd[:B] = 0 if seeds is None else np.asarray(seeds, np.uint64)
        self.seeds_host = sd
        self.seeds = torch.from_numpy(sd.view(np.int64)).to(dev)
        ti = np.zeros(bs, np.uint32)
        ti[:B] = np.arange(B) if trial_ids is None else np.asarray(trial_ids,
                                                                  np.uint32)
        self.trial_ids_host = ti
        self.trial_ids = torch.from_numpy(ti.view(np.int32)).to(dev)
        # per-trial clock: origin 0, no end, until a slot is (re)started
        self.t_origin_host = np.zeros(bs, np.int64)
        self.t_end_host = np.full(bs, NO_END, np.int64)
        self.t_origin = torch.from_numpy(self.t_origin_host).to(dev)
        self.t_end = torch.from_numpy(self.t_end_host).to(dev)
        # pull road: the last end step of each 64-trial word (idle words
        # are skipped); None = every word live
        self.word_end = None
        self.spikes_out = (torch.zeros(nb, dtype=torch.float32, **z)
                           if spikes_out else None)
        self.counts = torch.zeros(nb, dtype=torch.int32, **z) if counts else None
        self.pop_counts = (torch.zeros(n, dtype=torch.int32, **z)
                           if pop_counts else None)
        self.raster_len = int(raster_len)
        if self.raster_len and self.raster_len < L.MAX_WINDOW:
            raise ValueError("raster_len must be 0 or >= %d" % L.MAX_WINDOW)
        self.raster = (torch.zeros(self.raster_len * self.n_blocks
                                   * L.LINE_WORDS, dtype=torch.int64, **z)
                       if raster_len else None)
        if self.road == "pull":
            order, self.in_offsets, self.in_cols, self.in_vals = \
                net.pull_layout(bs)
            # the identity order goes to the kernel as a null pointer: no
            # dependent load of each tile's neuron numbers
            identity = not (net.locality and net.conn.groups is not None)
            self.order = None if identity else order
            self.rec = torch.zeros(L.REC_SLOTS * n * (bs // 64),
                                   dtype=torch.int64, **z)
            self.act = torch.zeros(n * max(1, bs // self.pull_tile),
                                   dtype=torch.int64, **z)
            # the records again as short trial lists, 16 B per (neuron,
            # tile part, slot); zero = empty lists, matching zero words
            self.rec_ids = torch.zeros(n * max(1, bs // self.pull_tile)
                                       * L.REC_SLOTS * 16,
                                       dtype=torch.uint8, **z)
            parts = max(1, bs // self.pull_tile)
            self.due = (torch.zeros(n * parts, dtype=torch.int64, **z)
                        if gate else None)
            # calm (neuron, part)s skip windows unread; any host write of
            # state or flags zeroes it (_calm_reset)
            self.calm = (torch.zeros((n * parts + 3) // 4 * 4,
                                     dtype=torch.uint8, **z)[:n * parts]
                         if gate else None)
            # a gated window's work list: two counters and two parities,
            # then the tiles it runs (zero at the start)
            per = max(1, self.pull_tile // bs)
            tiles = (n + per - 1) // per * parts
            self.work = (torch.zeros(4 + tiles, dtype=torch.int32, **z)
                         if gate else None)
        else:
            self.order = self.in_offsets = self.in_cols = self.in_vals = None
            self.rec = self.act = self.due = self.calm = None
            self.work = None
            self.rec_ids = None
        self.gate = self.due is not None
        self.t = 0  # host mirror of the device clock
        self._open_window = 0
        self.ring_until = 0  # pull road: steps < this also read the ring

        # Host mirrors of the host-written per-neuron-trial arrays.  Changes
        # are made here and sent down as one contiguous copy: scattered
        # index writes by many workgroups could share 128-byte lines.
        self.fl_host = np.zeros((n, bs), np.uint8)
        if bs > B:
            self.fl_host[:, B:] = L.FIRE_OFF
        self.vhold_host = np.zeros((n, bs), np.float32)
        self.version = 0  # bumped whenever the kernel struct changes
        st0 = np.full((n, bs), int(self.c["refrac_steps"]), np.uint8)
        free = self._neurons(no_refractory)
        if free.size:
            self.fl_host[free, :B] |= L.NOREFRAC
            st0[free, :B] = 0
        self.st.copy_(torch.from_numpy(st0.reshape(-1)))
        self._push_flags()
        self._sync_struct()

    # ------------------------------------------------------------ struct
    def _sync_struct(self):
        c = self.c
        s = L.FisSim()
        s.n, s.batch, s.bstride, s.ev_cap = self.n, self.batch, self.bstride, \
            self.ev_cap
        s.offsets = self.net.offsets.data_ptr()
        s.cols = self.net.cols.data_ptr()
        s.vals = self.net.vals.data_ptr()
        for name in ("v", "g", "st", "fl", "ring", "sbits", "events",
                     "ev_count", "clock", "seeds", "trial_ids", "t_origin",
                     "t_end"):
            setattr(s, name, getattr(self, name).data_ptr())
        s.vhold = self.vhold.data_ptr() if self.vhold is not None else None
        for name in ("w_scale", "a_syn", "a_mem", "v_rest", "v_reset", "v_th",
                     "gain", "prob_scale"):
            setattr(s, name, float(c[name]))
        s.refrac_steps = int(c["refrac_steps"])
        for name in ("spikes_out", "counts", "pop_counts", "raster"):
            t = getattr(self, name)
            setattr(s, name, t.data_ptr() if t is not None else None)
        s.raster_len = self.raster_len
        for name in ("in_offsets", "in_cols", "in_vals", "rec", "act",
                     "order", "word_end", "drive_group", "drive_rate",
                     "due", "calm", "rec_ids", "work"):
            t = getattr(self, name, None)
            setattr(s, name, t.data_ptr() if t is not None else None)
        dims = getattr(self, "_drive_dims", (0, 0, 0))
        s.drive_n_groups, s.drive_n_seg, s.drive_seg_steps = dims
        s.pull_tile = self.pull_tile
        base = getattr(self, "_defer_base", None)
        if base is not None:  # inside deferred_clock
            s.clock_ahead, s.defer_tick = self.t - base, 1
        self.struct = s
        self._sp = ctypes.byref(s)
        self.version = getattr(self, "version", 0) + 1

    # ------------------------------------------------------------ flags
    def _trials(self, trials):
        if trials is None:
            return np.arange(self.batch)
        t = np.asarray(trials, dtype=np.int64).reshape(-1)
        if t.size and (t.min() < 0 or t.max() >= self.batch):
            raise IndexError("trial index outside [0, %d)" % self.batch)
        return t

    def _neurons(self, idx):
        a = np.asarray(idx if idx is not None else [], dtype=np.int64).reshape(-1)
        if a.size and (a.min() < 0 or a.max() >= self.n):
            raise IndexError("neuron index outside [0, %d)" % self.n)
        return a

    def _calm_reset(self):
        """The host wrote state or flags: no (neuron, part) may skip its
        next window unread (fis_sim.calm)."""
        if getattr(self, "calm", None) is not None:
            self.calm.zero_()

    def _push_flags(self):
        self._calm_reset()
        self.fl.copy_(torch.from_numpy(self.fl_host.reshape(-1)))

    def set_flag(self, bit, neurons, trials=None):
        """Clear ``bit`` on the given trials, then set it on ``neurons``."""
        tr = self._trials(trials)
        ne = self._neurons(neurons)
        f = self.fl_host
        old = f.copy() if bit & L.OUT_OFF else None
        f[:, tr] &= np.uint8(0xFF ^ bit)
        if ne.size:
            f[np.ix_(ne, tr)] |= np.uint8(bit)
        self._push_flags()
        if old is not None:
            self._resend(old)

    def set_flag_pairs(self, bit, neurons, trials):
        """Set ``bit`` on explicit (neuron, trial) pairs (other entries kept)."""
        ne = self._neurons(neurons)
        tr = np.asarray(trials, dtype=np.int64).reshape(-1)
        if ne.size != tr.size:
            raise ValueError("neurons and trials differ in length")
        if tr.size and (tr.min() < 0 or tr.max() >= self.batch):
            raise IndexError("trial index outside [0, %d)" % self.batch)
        if ne.size:
            old = self.fl_host.copy() if bit & L.OUT_OFF else None
            self.fl_host[ne, tr] |= np.uint8(bit)
            self._push_flags()
            if old is not None:
                self._resend(old)

    def _resend(self, old):
        """The last step's spikes were sent with the old transmit flag; the
        reference masks them with the flag current at the next step
        (sim.py:910-918).  Correct where they went: the ring slot (push
        road, or the pull road right after a restore, when they were seeded
        into the ring) or the last step's spike record (pull road)."""
        changed = np.flatnonzero(((old ^ self.fl_host) & L.OUT_OFF).reshape(-1))
        if changed.size == 0:
            return
        idx = torch.from_numpy(changed.astype(np.int32)).to(self.device)
        if self.road == "pull" and self.t >= self.ring_until:
            L.check(self.lib.fis_pull_resend(self._sp, _ptr(idx),
                                             int(idx.numel()), _stream()),
                    "fis_pull_resend")
        else:
            L.check(self.lib.fis_resend(self._sp, _ptr(idx), int(idx.numel()),
                                        L.write_slot(self.t - 1), _stream()),
                    "fis_resend")
        self._resend_idx = idx  # kept alive until the launch has run

    def exempt_pairs(self, neurons, trials):
        """Refractory period 0 for (neuron, trial) pairs, before the first
        step: the counter starts at refrac_steps = 0 for them, as
        Engine.__init__ sets refrac = refrac_steps (sim.py:371-388)."""
        if self.t != 0:
            raise RuntimeError("refractory exemptions are set before stepping")
        ne = self._neurons(neurons)
        tr = np.asarray(trials, dtype=np.int64).reshape(-1)
        if ne.size:
            self.fl_host[ne, tr] |= np.uint8(L.NOREFRAC)
            self._push_flags()  # resets calm
            st = self.st.cpu().numpy().reshape(self.n, self.bstride)
            st[ne, tr] = 0
            self.st.copy_(torch.from_numpy(st.reshape(-1)))

    # ------------------------------------------------------------ drive
    def fuse_drive(self, table):
        """Push road: let the update launches draw ``table``'s kicks
        themselves (no separate drive launch per window; measured 1.20 ->
        1.06 us per step at B = 1).  ``None`` turns it off."""
        if table is None:
            self.drive_group = self.drive_rate = None
            self._sync_struct()
            return
        if self.road != "push":
            raise ValueError("the pull road draws a drive table in its own "
                             "launch (drive_table), not inside the window")
        if table.n_groups > 32767:
            raise ValueError("a fused drive table holds at most 32767 groups")
        g = np.full(self.nb, -1, np.int16)
        g[table.idx_host] = table.group_host.astype(np.int16)
        self.drive_group = torch.from_numpy(g).to(self.device)
        self.drive_rate = table.rate_hz
        self._drive_dims = (table.n_groups, table.n_seg, table.seg_steps)
        self._sync_struct()

    def fuse_drive_groups(self, groups, rate_hz, n_groups, n_seg=1,
                          seg_steps=1 << 30):
        """Push road: the update launches draw kicks for every
        neuron-trial idx with ``groups[idx] >= 0`` at rate
        ``rate_hz[min(t // seg_steps, n_seg - 1), groups[idx]]`` (t the
        trial's own step) -- the bits a drive table of those (idx, group)
        entries would set.  ``groups`` (int16 [n * bstride]) and
        ``rate_hz`` (float32 [n_seg * n_groups]) are device tensors the
        caller keeps and may change between windows."""
        if self.road != "push":
            raise ValueError("the pull road draws a drive table in its own "
                             "launch (drive_table), not inside the window")
        if not 0 < n_groups <= 32767 or n_seg < 1 or seg_steps < 1:
            raise ValueError("fused drive: 1..32767 groups, n_seg and "
                             "seg_steps >= 1")
        if (groups.dtype != torch.int16 or groups.numel() != self.nb
                or rate_hz.dtype != torch.float32
                or rate_hz.numel() != n_groups * n_seg
                or groups.device != self.device
                or rate_hz.device != self.device):
            raise ValueError("fused drive arrays: int16 [n * bstride] groups "
                             "and float32 [n_seg * n_groups] rates on %s"
                             % (self.device,))
        self.drive_group = groups
        self.drive_rate = rate_hz
        self._drive_dims = (int(n_groups), int(n_seg), int(seg_steps))
        self._sync_struct()

    # ------------------------------------------------------------ slots
    def restart_trials(self, trials, seeds, t_end, out_neurons=(),
                       out_trials=(), free_neurons=(), free_trials=()):
        """Start fresh trials in slots ``trials`` at the current clock.

        Their state is reset on the device (fis_reset_trials: the state
        Engine.__init__ gives, nothing in flight), their Philox key is
        ``seeds``, their own clock starts at 0 now and their spikes are
        counted until global step ``t_end``; (out_neurons, out_trials)
        pairs get m_out = 0 and (free_neurons, free_trials) pairs are
        refractory-exempt.  Other trials are not touched.
        """
        tr = self._trials(trials)
        if tr.size == 0:
            return
        self._calm_reset()
        sd = np.asarray(seeds, np.uint64).reshape(-1)
        te = np.asarray(t_end, np.int64).reshape(-1)
        if sd.size != tr.size or te.size != tr.size:
            raise ValueError("one seed and one end step per trial")
        dev = self.device
        tt = torch.from_numpy(tr.astype(np.int32)).to(dev)
        L.check(self.lib.fis_reset_trials(self._sp, _ptr(tt), int(tr.size),
                                          float(self.c["v0"]), _stream()),
                "fis_reset_trials")
        self.fl_host[:, tr] = 0
        self.vhold_host[:, tr] = 0.0
        keep = [tt]
        for bit, ne, tb in ((L.OUT_OFF, out_neurons, out_trials),
                            (L.NOREFRAC, free_neurons, free_trials)):
            ne = self._neurons(ne)
            tb = np.asarray(tb, np.int64).reshape(-1)
            if ne.size != tb.size:
                raise ValueError("neurons and trials differ in length")
            if ne.size == 0:
                continue
            if not np.isin(tb, tr).all():
                raise ValueError("flag pairs must lie in the restarted trials")
            self.fl_host[ne, tb] |= np.uint8(bit)
            idx = torch.from_numpy((ne * self.bstride + tb).astype(np.int32)
                                   ).to(dev)
            L.check(self.lib.fis_flag_pairs(self._sp, _ptr(idx),
                                            int(idx.numel()), bit, _stream()),
                    "fis_flag_pairs")
            keep.append(idx)
        self.seeds_host[tr] = sd
        self.seeds.copy_(torch.from_numpy(self.seeds_host.view(np.int64)))
        self.t_origin_host[tr] = self.t
        self.t_end_host[tr] = te
        self.t_origin.copy_(torch.from_numpy(self.t_origin_host))
        self.t_end.copy_(torch.from_numpy(self.t_end_host))
        self._push_word_end()
        self._slot_keep = keep  # index buffers live until the launches ran

    def set_trial_end(self, trials, t_end):
        """Stop counting (and, on the pull road, let a wholly idle 64-trial
        word be skipped) from global step ``t_end`` on."""
        tr = self._trials(trials)
        self.t_end_host[tr] = np.asarray(t_end, np.int64)
        self.t_end.copy_(torch.from_numpy(self.t_end_host))
        self._push_word_end()

    def track_idle_words(self):
        """Pull road: skip 64-trial words whose trials have all ended
        (t_end); for slot-based serving, where idle slots hold no result."""
        if self.road != "pull":
            return
        if self.word_end is None:
            self.word_end = torch.zeros(self.bstride // 64, dtype=torch.int64,
                                        device=self.device)
            self._sync_struct()
        self._push_word_end()

    def _push_word_end(self):
        if self.word_end is not None:
            we = self.t_end_host.reshape(-1, 64).max(axis=1)
            self.word_end.copy_(torch.from_numpy(np.ascontiguousarray(we)))

    def trial_state(self, trial):
        """One trial's column of the state at a window boundary, in a form
        any slot of any DeviceSim on the same connectome can load
        (:meth:`load_trial_state`): v, g, the refractory/spike byte, the
        flags and clamp values, the spike counts, the input in flight as
        ``inflight[j]`` = the integer synaptic input step ``t + j`` reads
        (j < 20: the spikes of steps t-20 .. t-1, the only ones still in
        flight), the trial's own step (``steps_done``), the steps it still
        counts (``steps_left``) and its Philox key.  The reference keeps
        the same quantities per engine (STATE_TENSORS, sim.py:775-846);
        here they are per trial, for moving a running trial between
        slots."""
        b = int(self._trials([trial])[0])
        n, bs = self.n, self.bstride

        def col(t):
            return t.view(n, bs)[:, b].clone()

        ring = self.ring_image().view(L.RING_LEN, n, bs)[:, :, b]
        slots = [L.read_slot(self.t + j) for j in range(L.DELAY)]
        return {
            "v": col(self.v), "g": col(self.g), "st": col(self.st),
            "counts": col(self.counts) if self.counts is not None else None,
            "inflight": ring[slots].clone(),
            "fl": self.fl_host[:, b].copy(),
            "vhold": self.vhold_host[:, b].copy(),
            "steps_done": int(self.t - self.t_origin_host[b]),
            "steps_left": int(self.t_end_host[b] - self.t),
            "seed": int(self.seeds_host[b]),
            "trial_id": int(self.trial_ids_host[b]),
            "n": n,
        }

    def load_trial_state(self, trial, state):
        """Continue a trial saved by :meth:`trial_state` in slot ``trial``
        of this simulation, at this simulation's clock: the slot is reset
        (nothing of its previous occupant stays in flight), the columns are
        written, the input in flight goes into the ring slots steps t ..
        t+19 read (on the pull road those steps also consume the ring), and
        the trial's own clock continues where it stopped, so its Philox
        stream and rate schedule go on unchanged."""
        if int(state["n"]) != self.n:
            raise ValueError("state of a %d-neuron network, this one has %d"
                             % (state["n"], self.n))
        b = int(self._trials([trial])[0])
        n, bs = self.n, self.bstride
        self._calm_reset()
        tt = torch.tensor([b], dtype=torch.int32, device=self.device)
        L.check(self.lib.fis_reset_trials(self._sp, _ptr(tt), 1,
                                          float(self.c["v0"]), _stream()),
                "fis_reset_trials")
        self._slot_keep = [tt]
        for name in ("v", "g", "st"):
            getattr(self, name).view(n, bs)[:, b].copy_(state[name])
        if self.counts is not None and state["counts"] is not None:
            self.counts.view(n, bs)[:, b].copy_(state["counts"])
        ring = self.ring.view(L.RING_LEN, n, bs)
        for j in range(L.DELAY):
            ring[L.read_slot(self.t + j), :, b].copy_(state["inflight"][j])
        if self.road == "pull":
            self.ring_until = max(self.ring_until, self.t + L.DELAY)
        self.fl_host[:, b] = state["fl"]
        self.vhold_host[:, b] = state["vhold"]
        self._push_flags()
        self.vhold.copy_(torch.from_numpy(self.vhold_host.reshape(-1)))
        self.seeds_host[b] = np.uint64(state["seed"])
        self.seeds.copy_(torch.from_numpy(self.seeds_host.view(np.int64)))
        self.trial_ids_host[b] = np.uint32(state["trial_id"])
        self.trial_ids.copy_(torch.from_numpy(
            self.trial_ids_host.view(np.int32)))
        self.t_origin_host[b] = self.t - int(state["steps_done"])
        self.t_end_host[b] = self.t + int(state["steps_left"])
        self.t_origin.copy_(torch.from_numpy(self.t_origin_host))
        self.t_end.copy_(torch.from_numpy(self.t_end_host))
        self._push_word_end()

    def trial_counts(self, trial):
        """Spike counts of one trial, int64 [n] (one column, on the host)."""
        if self.counts is None:
            raise RuntimeError("built without counts")
        col = self.counts.view(self.n, self.bstride)[:, int(trial)]
        return col.to(torch.int64).cpu().numpy()

    def trial_counts_async(self, trial):
        """Spike counts of one trial copied to pinned host memory on the
        current stream without waiting: (host int64 [n] tensor, event
        recorded after the copy).  The counts are there once the event has
        completed."""
        if self.counts is None:
            raise RuntimeError("built without counts")
        col = self.counts.view(self.n, self.bstride)[:, int(trial)]
        host = torch.empty(self.n, dtype=torch.int64, pin_memory=True)
        host.copy_(col.to(torch.int64), non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        return host, ev

    def set_hold(self, neurons, v_clamp, trials=None):
        tr = self._trials(trials)
        ne = self._neurons(neurons)
        self.vhold_host[:, tr] = 0.0
        if ne.size:
            self.vhold_host[np.ix_(ne, tr)] = np.float32(v_clamp)
        self.vhold.copy_(torch.from_numpy(self.vhold_host.reshape(-1)))
        self.set_flag(L.HOLD, neurons, trials)

    def write_trials(self, buf, arr):
        """Write [batch, n] values into a trial-minor buffer as one copy."""
        self._calm_reset()
        host = buf.cpu().numpy().reshape(self.n, self.bstride).copy()
        host[:, :self.batch] = np.asarray(arr).reshape(self.batch, self.n).T
        buf.copy_(torch.from_numpy(host.reshape(-1)))

    # ------------------------------------------------------------ launches
    def _check_rates(self, rates):
        if rates.dtype != torch.float32 or not rates.is_contiguous() or \
                tuple(rates.shape) != (self.batch, self.n) or \
                rates.device != self.device:
            raise ValueError("rates must be contiguous float32 [%d, %d] on %s"
                             % (self.batch, self.n, self.device))

    def drive_rates(self, rates, W=1):
        """Kicks for W steps from dense rates [batch, n] (Hz) on the device."""
        self._check_rates(rates)
        L.check(self.lib.fis_drive_rates(self._sp, _ptr(rates), W, _stream()),
                "fis_drive_rates")

    def step_rates(self, rates, W=1, checked=False):
        """One window driven by dense rates [batch, n] (Hz).  On the push
        road the update launch draws the kicks itself (fis_sim.drive_dense:
        one launch fewer per window, the bits fis_drive_rates would set);
        on the pull road, or with a fused drive table, fis_drive_rates runs
        first.  ``checked``: the caller has checked ``rates`` (contiguous
        float32 [batch, n] on this device)."""
        if self.road == "pull" or getattr(self, "drive_group", None) is not None:
            self.drive_rates(rates, W)
            self.step_launch(W)
            return
        if not checked:
            self._check_rates(rates)
        self.struct.drive_dense = rates.data_ptr()
        try:
            self.step_launch(W)
        finally:
            self.struct.drive_dense = None

    def drive_table(self, table, W):
        L.check(self.lib.fis_drive_table(
            self._sp, _ptr(table.entries), table.n_entries,
            _ptr(table.rate_hz), table.n_groups, table.n_seg, table.seg_steps,
            W, _stream()), "fis_drive_table")

    def drive_tape(self, tape, W):
        L.check(self.lib.fis_drive_tape(self._sp, _ptr(tape.ptr),
                                        _ptr(tape.idx), tape.n_steps, W,
                                        _stream()), "fis_drive_tape")

    def update(self, W):
        """W update steps; the window closes with :meth:`scatter`."""
        self._push_only("update")
        L.check(self.lib.fis_update(self._sp, W, _stream()), "fis_update")
        self._open_window = W

    def scatter(self):
        """Scatter the window's spikes and advance the clock by W."""
        self._push_only("scatter")
        W = self._open_window
        L.check(self.lib.fis_scatter(self._sp, W, _stream()), "fis_scatter")
        self._moved(W)

    def _push_only(self, what):
        if self.road != "push":
            raise RuntimeError("%s is a push-road launch; this simulation "
                               "runs the pull road (use window/pull)" % what)

    def pull(self, W):
        """One pull-road window of W steps: gather + update + clock += W."""
        ring_steps = max(0, min(W, self.ring_until - self.t))
        L.check(self.lib.fis_pull_window(self._sp, W, ring_steps, _stream()),
                "fis_pull_window")
        self._moved(W)

    def step_launch(self, W):
        """The road's launches of one window after its drive."""
        if self.road == "pull":
            self.pull(W)
        elif self.bstride < 64:
            self.update_scatter(W)
        else:
            self.update(W)
            self.scatter()

    def update_scatter(self, W):
        """Below 64 trials: the window's update and scatter as one launch
        (clock += W included)."""
        self._push_only("update_scatter")
        L.check(self.lib.fis_update_scatter(self._sp, W, _stream()),
                "fis_update_scatter")
        self._moved(W)

    def _moved(self, W):
        """The host clock moves on by a window of W steps.  In a deferred
        run the device clock stays put, and the next window's kernels read
        it plus the steps launched since the run began."""
        self.t += W
        base = getattr(self, "_defer_base", None)
        if base is not None:
            self.struct.clock_ahead = self.t - base

    @contextlib.contextmanager
    def deferred_clock(self):
        """Windows launched inside leave the device clock alone (each reads
        it plus the steps launched before it in the run) and one
        ``fis_tick`` at the end advances it by the whole run: for a graph
        of several windows, one clock launch per replay instead of one per
        window (at B = 1 that launch is a quarter of the window)."""
        if getattr(self, "_defer_base", None) is not None:
            raise RuntimeError("deferred_clock does not nest")
        self._defer_base = self.t
        self.struct.clock_ahead, self.struct.defer_tick = 0, 1
        try:
            yield self
        finally:
            steps = self.t - self._defer_base
            self._defer_base = None
            self.struct.clock_ahead, self.struct.defer_tick = 0, 0
            if steps:
                L.check(self.lib.fis_tick(self._sp, steps, _stream()),
                        "fis_tick")

    def window(self, W, drive=None):
        """One window: drive (a DriveTable, Tape or rates tensor), update,
        scatter."""
        if not 1 <= W <= L.MAX_WINDOW:
            raise ValueError("window of %d steps (1..%d)" % (W, L.MAX_WINDOW))
        if isinstance(drive, torch.Tensor):
            self.step_rates(drive, W)
            return
        if drive is not None:
            if isinstance(drive, DriveTable):
                self.drive_table(drive, W)
            elif isinstance(drive, Tape):
                self.drive_tape(drive, W)
            else:
                raise TypeError("unknown drive %r" % type(drive))
        self.step_launch(W)

    def read_kicks(self):
        """The Poisson kicks a drive launch has set for the window starting
        now (before the update consumes them), as (steps, neurons, trials)
        int64 arrays: the window's slice of a tape."""
        sb = self.sbits.view(self.n, self.bstride)[:, :self.batch] \
            .cpu().numpy().view(np.uint32)
        ne, tr = np.nonzero(sb)
        words = np.ascontiguousarray(sb[ne, tr])
        bits = np.unpackbits(words.view(np.uint8).reshape(-1, 4), axis=1,
                             bitorder="little")
        row, k = np.nonzero(bits)
        return ((self.t + k).astype(np.int64), ne[row].astype(np.int64),
                tr[row].astype(np.int64))

    def record_window(self, W, drive):
        """One window driven by ``drive`` (a DriveTable, or rates [batch, n]
        on the device), returning its kicks (:meth:`read_kicks`) -- a tape
        of the Philox drive that replays the same window exactly."""
        if isinstance(drive, torch.Tensor):
            self.drive_rates(drive, W)
        elif isinstance(drive, DriveTable):
            self.drive_table(drive, W)
        else:
            raise TypeError("record a DriveTable or a rates tensor")
        kicks = self.read_kicks()
        self.step_launch(W)
        return kicks

    def seed_ring(self, slot):
        L.check(self.lib.fis_seed_ring(self._sp, slot, _stream()),
                "fis_seed_ring")

    # ------------------------------------------------------------ views
    def view(self, t):
        """[batch, n] view of a trial-minor [n * bstride] tensor."""
        return t.view(self.n, self.bstride)[:, :self.batch].t()

    def refrac_counter(self):
        return self.view(self.st) & L.ST_RMAX

    def spike_state(self):
        return (self.view(self.st) & L.ST_SPIKE) != 0

    def ring_image(self):
        """The input in flight as the push road's int32 ring
        [40 * n * bstride] (spikes of step u in slot (u+21) mod 40): the
        ring itself on the push road; on the pull road the carried-over
        ring plus the spike records of the last 20 steps scattered into a
        copy (fis_pull_to_ring)."""
        if self.road == "push":
            return self.ring
        img = self.ring.clone()
        L.check(self.lib.fis_pull_to_ring(self._sp, _ptr(img), _stream()),
                "fis_pull_to_ring")
        return img

    def ring_counts(self, step_input):
        """Integer synaptic input that step ``step_input`` reads, i.e.
        the sum computed from the spikes of step ``step_input - 20``;
        [batch, n]."""
        slot = L.read_slot(step_input)
        return self.view(self.ring_image()[slot * self.nb:(slot + 1) * self.nb])

    def read_raster(self, t_from, t_to):
        """Spikes of steps [t_from, t_to) from the raster ring, bool
        [steps, batch, n].  Rows older than raster_len steps are gone."""
        if self.raster is None:
            raise RuntimeError("built without a raster")
        if t_to - t_from > self.raster_len or t_from < self.t - self.raster_len:
            raise ValueError("raster rows [%d, %d) no longer held"
                             % (t_from, t_to))
        rows = np.array([t % self.raster_len for t in range(t_from, t_to)])
        full = self.raster.cpu().numpy().reshape(
            self.raster_len, self.n_blocks, L.LINE_WORDS)
        words = full[rows][:, :, :L.BLOCK_WAVES].reshape(len(rows), -1)
        words = np.ascontiguousarray(words[:, :self.n_waves])
        bits = np.unpackbits(words.view(np.uint8), axis=1, bitorder="little")
        bits = bits[:, :self.nb].reshape(len(rows), self.n, self.bstride)
        return bits[:, :, :self.batch].transpose(0, 2, 1).astype(bool)

    def trial_raster(self, trial, t_from, t_to):
        """Spikes of one trial at steps [t_from, t_to), bool [steps, n],
        picked out of the raster ring on the device (only that trial's bits
        cross to the host)."""
        if self.raster is None:
            raise RuntimeError("built without a raster")
        if t_to - t_from > self.raster_len or t_from < self.t - self.raster_len:
            raise ValueError("raster rows [%d, %d) no longer held"
                             % (t_from, t_to))
        b = int(self._trials([trial])[0])
        dev = self.device
        idx = torch.arange(self.n, device=dev, dtype=torch.int64) \
            * self.bstride + b
        wave = idx >> 6
        col = (wave >> 2) * L.LINE_WORDS + (wave & 3)
        rows = torch.tensor([t % self.raster_len for t in range(t_from, t_to)],
                            device=dev, dtype=torch.int64)
        words = self.raster.view(self.raster_len, -1)[rows][:, col]
        bits = torch.bitwise_right_shift(words, (idx & 63)[None, :]) & 1
        return bits.to(torch.bool).cpu().numpy()

    def rec_slot(self, u):
        """The reference's rec(u) (sum of spikes of step u-1), int [B, n]."""
        slot = L.write_slot(u - 1)
        return self.view(self.ring_image()[slot * self.nb:(slot + 1) * self.nb])


class DriveTable:
    """Poisson-driven neuron-trials with a piecewise-constant rate schedule.

    ``idx`` are neuron-trial indices (neuron * bstride + trial), ``group``
    picks a column of ``rate_hz[n_seg, n_groups]``; segment = step //
    seg_steps (the last segment holds afterwards).
    """

    def __init__(self, sim: DeviceSim, neurons, trials, groups, rate_hz,
                 seg_steps=1 << 30):
        ne = np.asarray(neurons, dtype=np.int64).reshape(-1)
        tr = np.asarray(trials, dtype=np.int64).reshape(-1)
        gr = np.asarray(groups, dtype=np.int64).reshape(-1)
        rt = np.asarray(rate_hz, dtype=np.float32)
        if rt.ndim == 1:
            rt = rt[None, :]
        if not (ne.size == tr.size == gr.size):
            raise ValueError("neurons, trials and groups differ in length")
        if ne.size:
            if ne.min() < 0 or ne.max() >= sim.n:
                raise IndexError("driven neuron outside [0, %d)" % sim.n)
            if tr.min() < 0 or tr.max() >= sim.batch:
                raise IndexError("driven trial outside [0, %d)" % sim.batch)
            if gr.min() < 0 or gr.max() >= rt.shape[1]:
                raise IndexError("rate group outside [0, %d)" % rt.shape[1])
        idx = ne * sim.bstride + tr
        if np.unique(idx).size != idx.size:
            raise ValueError("a neuron-trial appears twice in the drive table")
        check_rates_hz(rt, "drive table")
        # the device copy in neuron-trial order: a wave's entries then set
        # kick words side by side (a few lines per wave) instead of one line
        # each (the entries are independent, so their order is free)
        order = np.argsort(idx, kind="stable")
        ent = np.empty((idx.size, 2), np.uint32)
        ent[:, 0] = idx[order]
        ent[:, 1] = gr[order]
        self.idx_host, self.group_host = idx, gr
        self.rate_host = np.ascontiguousarray(rt)
        self.n_entries = int(idx.size)
        self.n_seg, self.n_groups = int(rt.shape[0]), int(rt.shape[1])
        self.seg_steps = int(min(seg_steps, 1 << 30))
        dev = sim.device
        self.entries = torch.from_numpy(ent.view(np.int32).reshape(-1)).to(dev) \
            if idx.size else torch.zeros(2, dtype=torch.int32, device=dev)
        self.rate_hz = torch.from_numpy(np.ascontiguousarray(rt).reshape(-1)).to(dev)

    def set_rates(self, rate_hz):
        rt = np.asarray(rate_hz, dtype=np.float32).reshape(self.n_seg,
                                                           self.n_groups)
        check_rates_hz(rt, "drive table")
        self.rate_host = np.ascontiguousarray(rt)
        self.rate_hz.copy_(torch.from_numpy(rt.reshape(-1)))


TAPE_FORMAT = "flybrain_amd.tape/1"
TRIAL_STATE_FORMAT = "flybrain_amd.trial_state/1"


def save_trial_state(path, state):
    """Write a :meth:`DeviceSim.trial_state` dict as ``.npz`` (no pickled
    objects), so an evicted trial can be resumed by another process."""
    arrays = {"format": np.array(TRIAL_STATE_FORMAT)}
    for k, v in state.items():
        if isinstance(v, torch.Tensor):
            arrays[k] = v.cpu().numpy()
        elif v is None:
            continue
        else:
            arrays[k] = np.asarray(v)
    with open(path, "wb") as f:
        np.savez_compressed(f, **arrays)


def load_trial_state(path):
    """Read a trial state written by :func:`save_trial_state` (tensors on
    the host; :meth:`DeviceSim.load_trial_state` copies them over)."""
    with np.load(path, allow_pickle=False) as z:
        if "format" not in z.files or str(z["format"]) != TRIAL_STATE_FORMAT:
            raise ValueError("%s is not a %s file" % (path,
                                                      TRIAL_STATE_FORMAT))
        out = {"counts": None}
        for k in z.files:
            if k == "format":
                continue
            a = z[k]
            if k in ("fl", "vhold"):
                out[k] = a.copy()
            elif a.ndim == 0:
                out[k] = int(a)
            else:
                out[k] = torch.from_numpy(a.copy())
        return out


def save_tape(path, steps, neurons, trials, n, batch, seeds=None):
    """Write a Poisson tape as ``.npz`` (no pickled objects): the kicked
    (step, neuron, trial) triples, the shape [n_steps, batch, n] and,
    when the tape was recorded from the Philox drive, the trials' seeds.
    The reference's drive is Bernoulli(rate * DT/1000) per neuron-trial
    and step (lab/flybrain/sim.py:848-861); a tape is its outcomes, so a
    run fed the tape is reproducible by any engine of the same step --
    the reference's own included."""
    st = np.asarray(steps, np.int64).reshape(-1)
    ne = np.asarray(neurons, np.int64).reshape(-1)
    tr = np.asarray(trials, np.int64).reshape(-1)
    if not (st.size == ne.size == tr.size):
        raise ValueError("steps, neurons and trials differ in length")
    if ne.size and (ne.min() < 0 or ne.max() >= n or tr.min() < 0
                    or tr.max() >= batch or st.min() < 0):
        raise IndexError("tape event outside [steps, %d trials, %d neurons]"
                         % (batch, n))
    n_steps = int(st.max()) + 1 if st.size else 0
    order = np.lexsort((ne, tr, st))
    arrays = {"format": np.array(TAPE_FORMAT),
              "step": st[order], "neuron": ne[order].astype(np.int32),
              "trial": tr[order].astype(np.int32),
              "shape": np.array([n_steps, batch, n], np.int64)}
    if seeds is not None:
        arrays["seed"] = np.asarray(seeds, np.uint64).reshape(-1)
    with open(path, "wb") as f:
        np.savez_compressed(f, **arrays)


def load_tape(path):
    """Read a tape written by :func:`save_tape`: a dict with ``step``,
    ``neuron``, ``trial`` (int64 arrays), ``n_steps``, ``batch``, ``n``
    and ``seed`` (or None)."""
    with np.load(path, allow_pickle=False) as z:
        if "format" not in z.files or str(z["format"]) != TAPE_FORMAT:
            raise ValueError("%s is not a %s file" % (path, TAPE_FORMAT))
        n_steps, batch, n = (int(x) for x in z["shape"])
        return {"step": z["step"].astype(np.int64),
                "neuron": z["neuron"].astype(np.int64),
                "trial": z["trial"].astype(np.int64),
                "n_steps": n_steps, "batch": batch, "n": n,
                "seed": z["seed"].copy() if "seed" in z.files else None}


class Tape:
    """Recorded Poisson outcomes: kicked neuron-trial indices per step."""

    @classmethod
    def load(cls, sim: DeviceSim, path):
        """A tape file (:func:`save_tape`) on ``sim``'s device; its shape
        must match the simulation's neurons and trials."""
        z = load_tape(path)
        if z["n"] != sim.n or z["batch"] != sim.batch:
            raise ValueError("tape of %d trials x %d neurons, simulation of "
                             "%d x %d" % (z["batch"], z["n"], sim.batch,
                                          sim.n))
        return cls(sim, z["step"], z["neuron"], z["trial"])

    def __init__(self, sim: DeviceSim, steps, neurons, trials=None):
        st = np.asarray(steps, dtype=np.int64).reshape(-1)
        ne = np.asarray(neurons, dtype=np.int64).reshape(-1)
        tr = np.zeros_like(ne) if trials is None else \
            np.asarray(trials, dtype=np.int64).reshape(-1)
        if ne.size and (ne.min() < 0 or ne.max() >= sim.n or tr.min() < 0
                        or tr.max() >= sim.batch or st.min() < 0):
            raise IndexError("tape event outside the simulation")
        self.n_steps = int(st.max()) + 1 if st.size else 0
        order = np.argsort(st, kind="stable")
        st, ne, tr = st[order], ne[order], tr[order]
        ptr = np.zeros(self.n_steps + 1, np.int64)
        np.cumsum(np.bincount(st, minlength=self.n_steps), out=ptr[1:])
        idx = (ne * sim.bstride + tr).astype(np.int32)
        dev = sim.device
        self.ptr = torch.from_numpy(ptr).to(dev)
        self.idx = torch.from_numpy(idx if idx.size else
                                    np.zeros(1, np.int32)).to(dev)


MAX_RATE_HZ = 10000.0  # 1000 / DT (lab/flybrain/sim.py:131)


class RateTooHigh(ValueError):
    """A rate outside [0, 1000/DT] Hz (lab/flybrain/sim.py:134-135)."""


def check_rates_hz(values, where):
    """Reject rates outside [0, 10,000] Hz before they reach a kernel.

    Same bounds and same loud failure as ``check_rates_hz``
    (lab/flybrain/sim.py:138-161): at DT = 0.1 ms a probability above one
    has no meaning, and the check belongs to whoever writes rates.
    """
    a = np.asarray(values, dtype=np.float64)
    if a.size == 0:
        return values
    worst = float(np.nanmax(a))
    lowest = float(np.nanmin(a))
    if not (lowest >= 0.0) or worst > MAX_RATE_HZ:
        raise RateTooHigh(
            "%s: drive rate %.4g Hz (lowest %.4g) outside [0, %.0f]; the "
            "Poisson input is Bernoulli(rate * DT/1000) and needs a "
            "probability in [0, 1]" % (where, worst, lowest, MAX_RATE_HZ))
    return values
