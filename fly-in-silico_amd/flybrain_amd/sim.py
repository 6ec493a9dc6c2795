"""Drop-in for the reference engine surface of lab/flybrain/sim.py.

Same names, arguments, return values and error behaviour as the reference's
``Engine`` (sim.py:338-953), ``build_engine`` (:956-987), ``run``
(:990-1117), ``Result`` (:258-321), ``Stimulus`` (:51-71), ``RunConfig``
(:238-255), ``check_rates_hz``/``RateTooHigh`` (:131-161) and
``save``/``load`` (:1120-1174).  The step itself runs in libfis.so
(HIP, gfx950): there is no CPU road and no silent fallback -- a missing
library or a CPU device raises ``KernelUnavailable``.

Differences a caller can observe, all documented where they arise:
  * ``refrac`` reads the one-byte counter, i.e. min(reference value, 127)
    (only ``refrac >= refrac_steps <= 126`` ever matters to the dynamics);
  * the Poisson drive is this library's Philox4x32-10 stream keyed by
    (seed, trial, neuron, step) instead of torch's generator, so a given
    seed gives different -- equally distributed -- kicks than the reference;
    with a recorded tape (``step_tape``) the two agree bit for bit;
  * ``spikes`` is an output copy of the spike state: writing into it does
    not change the next step (use ``restore`` to set state).
"""

from __future__ import annotations

import json
import os
import time
from dataclasses import asdict, dataclass, field
from pathlib import Path

import numpy as np
import torch

from . import _lib as L
from . import connectome as C
from .device import (MAX_RATE_HZ, DeviceNet, DeviceSim, DriveTable,  # noqa: F401
                     RateTooHigh, Tape, check_rates_hz, pick_road)
from .params import DT, MODEL_PARAMS

KernelUnavailable = L.KernelUnavailable
# The reference raises EventKernelUnavailable when its kernel cannot run
# (lab/flybrain/evspmv.py:173-190); callers that catch it keep working.
EventKernelUnavailable = L.KernelUnavailable

# Road names (sim.py:164-166).  This engine has one road, the HIP event
# scatter; the reference's two names are accepted because both compute the
# same exact product, and ``spmv_info`` reports what actually ran.
SPMV_CSR = "эталонный"
SPMV_EVENT = "событийный"
SPMV_HIP = "hip"
SPMV_CHOICES = (SPMV_HIP, SPMV_CSR, SPMV_EVENT)
# The default road comes from the environment like the reference's
# (sim.py:235); every accepted name runs the HIP road, and an unknown name
# is refused when an Engine is built.
SPMV_DEFAULT = os.environ.get("FLYBRAIN_SPMV", SPMV_HIP)

FAST = "быстрый"
REFERENCE = "эталонный"


@dataclass
class Stimulus:
    """Poisson drive into a set of neurons, optionally time-windowed."""

    indices: np.ndarray
    rate_hz: float = 200.0
    t_on_ms: float = 0.0
    t_off_ms: float | None = None
    name: str = "stim"

    def to_dict(self, with_indices=False):
        d = {"name": self.name, "n": int(len(self.indices)),
             "rate_hz": self.rate_hz, "t_on_ms": self.t_on_ms,
             "t_off_ms": self.t_off_ms}
        if with_indices:
            d["indices"] = [int(i) for i in self.indices]
        return d


@dataclass
class RunConfig:
    t_run_s: float = 1.0
    bin_ms: float = 10.0
    n_trials: int = 1
    seed: int = 0
    device: str | None = None
    params: dict = field(default_factory=lambda: dict(MODEL_PARAMS))
    speed: str = FAST
    spmv: str | None = None
    # extension: the engine's road ("auto", "push" or "pull"); results are
    # the same on either
    road: str = "auto"


class Result:
    """Per-neuron rates plus a neuron x time-bin raster (sim.py:258-321)."""

    def __init__(self, counts, bins, config, stimuli, silenced, neurons,
                 timing):
        self.counts = counts  # int64[N], spikes summed over trials
        self.bins = bins  # int16[n_bins, N], per bin summed over trials
        self.config = config
        self.stimuli = stimuli
        self.silenced = silenced
        self.neurons = neurons
        self.timing = timing

    @property
    def n_bins(self):
        return self.bins.shape[0]

    @property
    def rates(self):
        """Firing rate in Hz per neuron, averaged over trials."""
        return self.counts / (self.config.t_run_s * self.config.n_trials)

    def active(self, min_rate=0.5):
        return np.nonzero(self.rates >= min_rate)[0]

    # the reference's neuron-table columns of a rate table (sim.py:286-289)
    NEURON_COLUMNS = ("root_id", "label", "cell_class", "super_class",
                      "region", "super_region", "side", "top_nt")

    def rate_table(self, min_rate=0.5):
        """The active neurons as a DataFrame indexed by neuron, by rate
        descending, with ``rate_hz`` and ``n_spikes`` (sim.py:282-292).
        The FlyWire neuron-table columns come first where ``neurons`` is
        such a table; with a plain index array there are none."""
        import pandas as pd

        idx = self.active(min_rate)
        nt = self.neurons
        if isinstance(nt, pd.DataFrame) and all(
                c in nt.columns for c in self.NEURON_COLUMNS):
            df = nt.loc[idx, list(self.NEURON_COLUMNS)].copy()
        else:
            df = pd.DataFrame(index=pd.Index(idx))
        df["rate_hz"] = self.rates[idx]
        df["n_spikes"] = self.counts[idx]
        return df.sort_values("rate_hz", ascending=False)

    def region_timeline(self):
        """DataFrame region x bin of the mean firing rate in Hz
        (sim.py:294-306): bins summed over each region's neurons, divided
        by the region's size, the bin width and the trial count.  Regions
        come from ``neurons["region"]`` where ``neurons`` is a FlyWire
        neuron table; otherwise from the connectome's neuropil groups
        (``group_<k>``)."""
        import pandas as pd

        nt = self.neurons
        if isinstance(nt, pd.DataFrame) and "region" in nt.columns:
            reg = nt["region"].to_numpy()
        elif getattr(self, "groups", None) is not None:
            reg = np.array(["group_%d" % g for g in self.groups])
        else:
            raise ValueError("no regions: neurons is not a table with a "
                             "'region' column and the connectome has no "
                             "groups")
        codes, uniq = pd.factorize(reg)
        n_reg = len(uniq)
        acc = np.zeros((n_reg, self.n_bins), dtype=np.float64)
        np.add.at(acc, codes, self.bins.T.astype(np.float64))
        sizes = np.bincount(codes, minlength=n_reg).astype(np.float64)
        bin_s = self.config.bin_ms / 1000.0
        rates = acc / (sizes[:, None] * bin_s * self.config.n_trials)
        return pd.DataFrame(rates, index=list(uniq))

    def meta(self, with_indices=False):
        return {
            "config": {k: v for k, v in asdict(self.config).items()
                       if k != "params"},
            "params": self.config.params,
            "stimuli": [s.to_dict(with_indices) for s in self.stimuli],
            "silenced": ([int(i) for i in self.silenced] if with_indices
                         else None),
            "n_silenced": int(len(self.silenced)),
            "timing": self.timing,
            "n_neurons": int(len(self.neurons)),
            "n_active": int((self.rates >= 0.5).sum()),
            "total_spikes": int(self.counts.sum()),
        }


def _as_connectome(weights, n):
    if isinstance(weights, C.Connectome):
        conn = weights
    else:
        conn = C.from_reference_coo(weights)
    if conn.n != n:
        raise ValueError("weight matrix is %d x %d but n = %d"
                         % (conn.n, conn.n, n))
    return conn


def _silence(conn, silenced):
    """Drop every edge of silenced neurons (sim.py:324-335)."""
    return conn.silence(silenced)


class _Gen:
    """The drive stream's state: 64-bit seed and the step counter.

    Stands where the reference keeps ``torch.Generator`` (sim.py:358-359);
    get_state/set_state round-trip through a uint8 tensor like torch's.
    """

    def __init__(self, engine, seed):
        self._engine = engine
        self.seed = int(seed) & ((1 << 64) - 1)

    def get_state(self):
        a = np.array([self.seed, self._engine._sim.t], dtype=np.uint64)
        return torch.from_numpy(a.view(np.uint8).copy())

    def set_state(self, state):
        a = np.asarray(torch.as_tensor(state).cpu().numpy(),
                       dtype=np.uint8).view(np.uint64)
        self.seed = int(a[0])
        self._engine._set_clock(int(a[1]))
        self._engine._sim.seeds_host[:self._engine.batch] = self.seed
        self._engine._sim.seeds.copy_(torch.from_numpy(
            self._engine._sim.seeds_host.view(np.int64)))

    def manual_seed(self, seed):
        self.set_state(torch.from_numpy(np.array(
            [int(seed), self._engine._sim.t], np.uint64).view(np.uint8)))
        return self


class Engine:
    """One brain (``batch`` trials) held in HBM and stepped on the GPU."""

    STATE_TENSORS = ("v", "conductance", "spikes", "refrac", "delay_buffer")
    MASK_TENSORS = ("m_fire", "m_out", "m_in", "m_hold", "v_hold")

    def __init__(self, weights, n, params=None, device="cuda", batch=1,
                 seed=0, no_refractory=None, spmv=None, allow_inexact=False,
                 road="auto"):
        p = dict(params or MODEL_PARAMS)  # taken whole, as sim.py:350
        self.p = p
        self.n = int(n)
        self.batch = int(batch)
        dev = torch.device(device if device is not None else "cuda")
        if dev.type != "cuda":
            raise KernelUnavailable(
                "this engine steps on the GPU only (libfis.so, gfx950); "
                "device %r has no road here" % (device,))
        self.device = dev
        name = SPMV_DEFAULT if spmv is None else str(spmv)
        if name not in SPMV_CHOICES:
            raise ValueError("spmv road %r not in %s" % (name, SPMV_CHOICES))
        self._spmv_requested = name
        self.spmv = SPMV_HIP
        conn = _as_connectome(weights, self.n)
        ex = conn.exactness()
        self._spmv_exact = ex
        self._spmv_inexact_ok = bool(allow_inexact)
        if not ex["ok"] and not allow_inexact:
            raise KernelUnavailable(
                "the integer event sum is not equal to the reference's float32 "
                "csrmv on this matrix: largest per-neuron sum of |w| %.0f >= "
                "2^24.  Pass allow_inexact=True to run the exact integer sum "
                "anyway (it is then recorded as differing from csrmv)."
                % ex["max_row_abs_sum"])
        self.conn = conn
        self._net = DeviceNet(conn, dev)
        self.device = dev = self._net.device
        self.tau_syn_f = DT / p["tauSyn"]
        self.tau_mem_f = DT / p["tauMem"]
        self.prob_scale = DT / 1000.0
        self.poisson_gain = p["scalePoisson"]
        self.w_scale = p["wScale"]
        self.v_rest, self.v_reset = p["vRest"], p["vReset"]
        self.v_th = p["vThreshold"]
        self.buf_len = int(p["tDelay"] / DT) + 1
        free = None
        if no_refractory is not None and len(no_refractory):
            free = (no_refractory.cpu().numpy()
                    if torch.is_tensor(no_refractory)
                    else np.asarray(no_refractory, dtype=np.int64))
        self._free = free
        # road: "push" (event scatter into the int32 ring) or "pull"
        # (per-window gather from spike records, 64..1024 trials).  step()
        # advances one step per call, and the pull road's gather costs
        # nearly as much for one step as for a window of 20 (measured per
        # step at 138,639 neurons: B = 64 214 us pull against 124 us push,
        # B = 256 685 against 530), so "auto" is the push road here; run()
        # advances in windows of up to 20 and takes the pull road from 64
        # trials up.  Same results either way.
        self._sim = DeviceSim(self._net, self.batch, params=p,
                              seeds=[int(seed) & ((1 << 64) - 1)] * self.batch,
                              no_refractory=free, spikes_out=True,
                              road="push" if road == "auto" else road)
        self.road = self._sim.road
        # the [batch, n] view of the spike output, made once: step returns
        # this same tensor every call, updated in place (sim.py:745-771)
        self._spikes = self._sim.view(self._sim.spikes_out)
        self._clock_run = None  # an open deferred-clock run of eager steps
        self.gen = _Gen(self, seed)
        self.refrac_steps = torch.full((self.n,), float(round(p["tRefrac"] / DT)),
                                       dtype=torch.float32, device=dev)
        if free is not None:
            self.refrac_steps[torch.as_tensor(free, device=dev)] = 0.0
        self._v_clamp = self.v_reset
        self._rates = torch.zeros(self.batch, self.n, dtype=torch.float32,
                                  device=dev)
        self._graph = None
        self._graphs = None
        self.calls = 0

    # --- state views ----------------------------------------------------
    @property
    def v(self):
        return self._sim.view(self._sim.v)

    @property
    def conductance(self):
        return self._sim.view(self._sim.g)

    @property
    def spikes(self):
        return self._spikes

    @property
    def refrac(self):
        """Steps since the last spike, saturating at 127, as float32."""
        return self._sim.refrac_counter().to(torch.float32)

    @property
    def ptr(self):
        return self._sim.t % self.buf_len

    @property
    def delay_buffer(self):
        """The reference's [B, 19, n] ring of synaptic input in mV
        (assembled on the host from one copy of the int32 ring)."""
        s = self._sim
        T = s.t
        ring = s.ring_image().cpu().numpy().reshape(L.RING_LEN, s.n,
                                                    s.bstride)
        w = np.float32(self.w_scale)
        out = np.zeros((self.batch, self.buf_len, self.n), np.float32)
        for u in range(max(0, T - self.buf_len), T):
            I = ring[L.write_slot(u - 1), :, :self.batch].T
            out[:, u % self.buf_len, :] = w * I.astype(np.float32)
        return torch.from_numpy(out)

    def _mask(self, bit):
        f = self._sim.view(self._sim.fl)
        return ((f & bit) == 0).to(torch.float32)

    @property
    def m_fire(self):
        return self._mask(L.FIRE_OFF)

    @property
    def m_out(self):
        return self._mask(L.OUT_OFF)

    @property
    def m_in(self):
        return self._mask(L.IN_OFF)

    @property
    def m_hold(self):
        return self._mask(L.HOLD)

    @property
    def v_hold(self):
        return self._sim.view(self._sim.vhold) * (1.0 - self.m_hold)

    # --- road report ------------------------------------------------------
    def spmv_info(self):
        road = ("libfis event scatter, int32 ring (gfx950)"
                if self.road == "push" else
                "libfis windowed gather from spike records (gfx950)")
        out = {"умножение связей": self.spmv,
               "road": road,
               "requested": self._spmv_requested,
               "побитовое равенство обеспечено": bool(self._spmv_exact["ok"]),
               "условия равенства": self._spmv_exact}
        if self._spmv_inexact_ok:
            out["разрешено приближение"] = True
        return out

    # --- interventions (sim.py:510-629) ------------------------------------
    def set_masks(self, fire=None, out=None, inc=None, hold=None,
                  v_clamp=None):
        """Intervention masks by neuron index lists, all trials.

        ``fire``: no spike output (v keeps integrating); ``out``: spike
        happens but is not transmitted (19-step tail); ``inc``: nothing
        received, instantly; ``hold``: v replaced by ``v_clamp`` (default
        vReset) before the threshold.  An empty list lifts that mask.
        """
        s = self._sim
        if fire is not None:
            s.set_flag(L.FIRE_OFF, fire)
        if out is not None:
            s.set_flag(L.OUT_OFF, out)
        if inc is not None:
            s.set_flag(L.IN_OFF, inc)
        if hold is not None:
            self._v_clamp = self.v_reset if v_clamp is None else float(v_clamp)
            s.set_hold(hold, self._v_clamp)

    def fan(self):
        got = getattr(self, "_fan", None)
        if got is None:
            n_in, n_out = self.conn.fan()
            got = self._fan = (torch.from_numpy(n_in), torch.from_numpy(n_out))
        return got

    def mask_counts(self):
        f = self._sim.view(self._sim.fl)[0]
        return {"не выдают спайк": int(((f & L.FIRE_OFF) != 0).sum()),
                "исходящие отключены": int(((f & L.OUT_OFF) != 0).sum()),
                "входящие отключены": int(((f & L.IN_OFF) != 0).sum()),
                "мембрана удержана": int(((f & L.HOLD) != 0).sum())}

    def masks_clear(self):
        f = self._sim.view(self._sim.fl)
        return not bool(((f & (L.FIRE_OFF | L.OUT_OFF | L.IN_OFF | L.HOLD))
                         != 0).any())

    def advance_ptr(self):
        """Advance the ring position without a step (sim.py:630-631)."""
        self._set_clock(self._sim.t + 1)

    def _set_clock(self, t):
        self._clock_flush()
        s = self._sim
        s.t = int(t)
        s.clock.fill_(int(t))

    # Eager steps run as one long deferred-clock run: each launch carries the
    # steps taken since the device clock last moved (clock_ahead), and the
    # clock catches up in one launch only when it is set, captured or a
    # graph replays -- one launch fewer per step (B = 1: 19.0 -> 16.4 us).
    def _clock_defer(self):
        s = self._sim
        if self._clock_run is not None and s.struct.clock_ahead >= (1 << 30):
            self._clock_flush()  # keep clock_ahead far inside int32
        if self._clock_run is None:
            self._clock_run = s.deferred_clock()
            self._clock_run.__enter__()

    def _clock_flush(self):
        run, self._clock_run = getattr(self, "_clock_run", None), None
        if run is not None:
            run.__exit__(None, None, None)

    # --- stepping -----------------------------------------------------------
    def _launch_step(self, rates):
        self._clock_defer()
        self._sim.step_rates(rates, 1, checked=True)

    def use_graphs(self, enable=True, force=False):
        """Capture one step (drive + update + scatter) into a HIP graph.

        Returns whether steps now replay a graph.  On MI355X a captured
        step replays slower than its direct launches (B = 1: 22 against 5
        us per step), so, as the reference stays ungraphed where graphs do
        not serve (sim.py:676-677, 715-737), this declines and returns False
        unless ``force`` -- results are the same either way.

        One graph serves every ring position: the step counter lives in
        device memory, so nothing in the graph is frozen per slot (the
        reference needs 19 graphs for that, sim.py:700-714).  A warm-up step
        runs first and the state is restored afterwards, as in the
        reference; on a capture failure the state is restored and the error
        re-raised.
        """
        if not enable or not force:
            self._graph = self._graphs = None
            return False
        if self._graph is not None:
            return True
        self._clock_flush()  # the graph's own step advances the clock
        snap = self._device_snapshot()
        try:
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                self._sim.step_rates(self._rates, 1)
            torch.cuda.current_stream().wait_stream(side)
            self._device_restore(snap)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._sim.step_rates(self._rates, 1)
            self._device_restore(snap)
        except BaseException:
            self._device_restore(snap)
            self._graph = self._graphs = None
            raise
        self._graph = g
        self._graphs = [g]
        self._graph_version = self._sim.version
        return True

    def _device_snapshot(self):
        s = self._sim
        names = ["v", "g", "st", "ring", "sbits", "ev_count", "clock",
                 "spikes_out"]
        names += [k for k in ("rec", "act", "due", "calm", "rec_ids")
                  if getattr(s, k, None) is not None]
        return ({k: getattr(s, k).clone() for k in names}, s.t)

    def _device_restore(self, snap):
        bufs, t = snap
        for k, val in bufs.items():
            getattr(self._sim, k).copy_(val)
        self._sim.t = t

    def step(self, rates):
        """Advance one 0.1 ms tick.  ``rates`` is [batch, n] in Hz."""
        self.calls += 1
        r = rates
        # (the cheapest tests first: this is the per-step host path)
        if not (type(r) is torch.Tensor and r.dtype is torch.float32
                and r.is_cuda and r.get_device() == self.device.index
                and r.is_contiguous()):
            r = torch.as_tensor(r, dtype=torch.float32).to(self.device).contiguous()
        if r.shape != (self.batch, self.n):
            raise ValueError("rates must be [%d, %d], got %s"
                             % (self.batch, self.n, tuple(r.shape)))
        if self._graph is not None and self._graph_version != self._sim.version:
            # the kernel arguments changed since capture: capture again
            self._graph = self._graphs = None
            self.use_graphs(force=True)
        # (on the pull road the 20 steps after a restore also read the
        # restored ring: launched directly, the graph does not carry that)
        if self._graph is not None and self._sim.t >= self._sim.ring_until:
            if r.data_ptr() != self._rates.data_ptr():
                self._rates.copy_(r)
            self._clock_flush()
            self._graph.replay()
            self._sim.t += 1
        else:
            self._launch_step(r)
        return self.spikes

    def step_tape(self, tape: Tape):
        """One step with the kicks of a recorded tape (parity road)."""
        self.calls += 1
        self._clock_defer()
        s = self._sim
        s.drive_tape(tape, 1)
        s.step_launch(1)
        return self.spikes

    def advance(self, steps, drive=None):
        """``steps`` steps in windows of up to 20 with one drive source
        (DriveTable, Tape or a [batch, n] rates tensor held constant)."""
        left = int(steps)
        self._clock_defer()
        while left > 0:
            W = min(L.MAX_WINDOW, left)
            self._sim.window(W, drive)
            left -= W
        self.calls += int(steps)

    # --- snapshot / restore (sim.py:775-846) ---------------------------------
    def snapshot(self):
        out = {}
        out["v"] = self.v.cpu().numpy().copy()
        out["conductance"] = self.conductance.cpu().numpy().copy()
        out["spikes"] = self._sim.spike_state().to(torch.float32).cpu().numpy()
        out["refrac"] = self.refrac.cpu().numpy().copy()
        out["delay_buffer"] = self.delay_buffer.cpu().numpy()
        for name in self.MASK_TENSORS:
            out[name] = getattr(self, name).cpu().numpy().copy()
        out["ptr"] = np.int64(self.ptr)
        out["gen"] = self.gen.get_state().numpy().copy()
        out["calls"] = np.int64(self.calls)
        return out

    def restore(self, state, drop=()):
        """Put a snapshot back in place; returns the names restored.

        Buffers are written in place, so a captured graph stays valid.  The
        step counter comes from ``gen`` (this engine's drive state) when
        present, else from ``calls``; it must agree with ``ptr``.
        """
        s = self._sim
        s._calm_reset()  # state is written from the host below
        done = []
        B, n, bs = self.batch, self.n, s.bstride
        T = s.t
        if "gen" not in drop and "gen" in state:
            a = np.asarray(state["gen"], dtype=np.uint8).view(np.uint64)
            T = int(a[1])
            self.gen.seed = int(a[0])
            s.seeds_host[:B] = self.gen.seed
            s.seeds.copy_(torch.from_numpy(s.seeds_host.view(np.int64)))
            done.append("gen")
        elif "calls" in state:
            T = int(state["calls"])
        if "ptr" not in drop and "ptr" in state:
            if T % self.buf_len != int(state["ptr"]) % self.buf_len:
                raise ValueError("snapshot ptr %d disagrees with step %d"
                                 % (int(state["ptr"]), T))
            done.append("ptr")

        if "v" not in drop and "v" in state:
            s.write_trials(s.v, np.asarray(state["v"], np.float32))
            done.append("v")
        if "conductance" not in drop and "conductance" in state:
            s.write_trials(s.g, np.asarray(state["conductance"], np.float32))
            done.append("conductance")
        st = s.st.cpu().numpy().reshape(n, bs).copy()
        if "refrac" not in drop and "refrac" in state:
            r = np.minimum(np.asarray(state["refrac"], np.float64),
                           L.ST_RMAX).astype(np.uint8).reshape(B, n).T
            st[:, :B] = (st[:, :B] & L.ST_SPIKE) | r
            done.append("refrac")
        if "spikes" not in drop and "spikes" in state:
            sp = (np.asarray(state["spikes"]).reshape(B, n).T != 0)
            st[:, :B] = (st[:, :B] & L.ST_RMAX) | (sp.astype(np.uint8)
                                                  * L.ST_SPIKE)
            done.append("spikes")
        s.st.copy_(torch.from_numpy(st.reshape(-1)))
        masks_changed = False
        for name, bit in (("m_fire", L.FIRE_OFF), ("m_out", L.OUT_OFF),
                          ("m_in", L.IN_OFF), ("m_hold", L.HOLD)):
            if name in drop or name not in state:
                continue
            off = (np.asarray(state[name]).reshape(B, n).T == 0)
            f = s.fl_host[:, :B]
            f &= np.uint8(0xFF ^ bit)
            f |= off.astype(np.uint8) * np.uint8(bit)
            masks_changed = True
            done.append(name)
        if masks_changed:
            s._push_flags()
        if "v_hold" not in drop and "v_hold" in state:
            vh = np.asarray(state["v_hold"], np.float32).reshape(B, n)
            s.vhold_host[:, :B] = vh.T
            s.vhold.copy_(torch.from_numpy(s.vhold_host.reshape(-1)))
            done.append("v_hold")
        self._set_clock(T)
        if "delay_buffer" not in drop and "delay_buffer" in state:
            self._restore_ring(np.asarray(state["delay_buffer"], np.float32), T)
            done.append("delay_buffer")
        if s.spikes_out is not None:
            s.spikes_out.copy_(torch.from_numpy(
                ((st & L.ST_SPIKE) != 0).astype(np.float32).reshape(-1)))
        if "calls" in state:
            self.calls = int(state["calls"])
        return done

    def _restore_ring(self, db, T):
        """Rebuild the int32 ring from the reference-format ring + spikes.

        rec(u) for u in T-19..T-1 comes from the snapshot (exact inverse of
        rec = 0.275f * I, checked); rec(T) = W @ (spikes * m_out) is
        recomputed by scattering the current spikes.
        """
        s = self._sim
        w = np.float32(self.w_scale)
        ring = np.zeros((L.RING_LEN, s.n, s.bstride), np.int32)
        for u in range(T - self.buf_len, T):
            if u < 0:
                continue
            rec = db[:, u % self.buf_len, :]
            I = np.rint(rec.astype(np.float64) / float(w)).astype(np.int64)
            if not np.array_equal((w * I.astype(np.float32)).astype(np.float32),
                                  rec + np.float32(0.0)):
                raise ValueError("delay_buffer slot %d is not a multiple of "
                                 "wScale: not a state of this model"
                                 % (u % self.buf_len))
            ring[L.write_slot(u - 1), :, :self.batch] = I.T.astype(np.int32)
        s.ring.copy_(torch.from_numpy(ring.reshape(-1)))
        s.seed_ring(L.write_slot(T - 1))
        if s.road == "pull":
            # everything in flight is now in the int32 ring, which the next
            # 20 steps consume; the spike records hold nothing older
            s.rec.zero_()
            s.rec_ids.zero_()
            s.act.zero_()
            s.ring_until = T + L.DELAY


def build_engine(weights=None, neurons=None, silenced=None, params=None,
                 device=None, batch=1, seed=0, no_refractory=None,
                 graphs=True, spmv=None, allow_inexact=False, road="auto"):
    """Load what is needed and return a ready :class:`Engine`
    (sim.py:956-987).  ``weights``: reference-format sparse COO W[post, pre],
    a ``{indices, values, shape}`` dict, or a Connectome."""
    if weights is None:
        # the reference's default: its weight cache, else the FlyWire
        # tables it reads (data.py:221-259); raises naming both paths
        weights = C.load_default()
    if neurons is None:
        n = weights.n if isinstance(weights, C.Connectome) else \
            int((weights["shape"] if isinstance(weights, dict)
                 else weights.shape)[0])
    else:
        n = len(neurons)
    device = device or "cuda"
    conn = _as_connectome(weights, n)
    conn = _silence(conn, np.asarray(silenced if silenced is not None else [],
                                     dtype=np.int64))
    eng = Engine(conn, n, params=params, device=device, batch=batch, seed=seed,
                 no_refractory=no_refractory, spmv=spmv,
                 allow_inexact=allow_inexact, road=road)
    if graphs:
        eng.use_graphs()
    return eng


def schedule(stimuli, steps):
    """Rate-change events (step, indices, rate) of a stimulus list, sorted
    by step with ties in stimulus order (sim.py:1021-1031)."""
    events = []
    for s in stimuli:
        idx = np.asarray(s.indices, dtype=np.int64)
        check_rates_hz(s.rate_hz, "stimulus %r" % s.name)
        on = int(round(s.t_on_ms / DT))
        off = steps if s.t_off_ms is None else int(round(s.t_off_ms / DT))
        events.append((on, idx, float(s.rate_hz)))
        if off < steps:
            events.append((off, idx, 0.0))
    events.sort(key=lambda e: e[0])
    return events


def rates_trace(stimuli, n, steps):
    """Per-step rate vector the schedule produces, float32 [steps, n]
    (what the reference writes into ``rates`` before each step)."""
    ev = schedule(stimuli, steps)
    rates = np.zeros(n, np.float32)
    out = np.zeros((steps, n), np.float32)
    k = 0
    for t in range(steps):
        while k < len(ev) and ev[k][0] == t:
            rates[ev[k][1]] = ev[k][2]
            k += 1
        out[t] = rates
    return out


def run(stimuli, silenced=None, config: RunConfig | None = None, neurons=None,
        weights=None, progress=True, on_bin=None, should_stop=None,
        tape=None):
    """Simulate and return a :class:`Result` (sim.py:990-1117).

    The schedule, bins and counts follow the reference: rates switch at
    round(t_on/DT) and round(t_off/DT), later events at the same step win,
    every stimulated neuron is exempt from refractoriness, bins are
    per-neuron spike counts summed over trials as int16, ``on_bin`` gets
    each closed bin and ``should_stop`` is polled between bins.

    Steps run in windows of up to 20 that never cross a bin edge or a
    schedule change, with the drive read from a device table.

    ``tape`` (extension): (steps, neurons, trials) arrays of recorded
    Bernoulli outcomes, or the path of a tape file
    (``device.save_tape``); when given they replace the Philox draw, so
    the run can be checked against the reference fed the same tape.
    """
    config = config or RunConfig()
    silenced = np.asarray([] if silenced is None else silenced, dtype=np.int64)
    if weights is None:
        weights = C.load_default()  # as build_engine (data.py:221-259)
    if neurons is None:
        n = weights.n if isinstance(weights, C.Connectome) else \
            int((weights["shape"] if isinstance(weights, dict)
                 else weights.shape)[0])
        neurons = np.arange(n)
    device = config.device or "cuda"
    p = config.params
    n = len(neurons)
    t0 = time.perf_counter()
    B = config.n_trials
    steps = int(round(config.t_run_s * 1000.0 / DT))
    bin_steps = max(1, int(round(config.bin_ms / DT)))
    n_bins = int(np.ceil(steps / bin_steps))

    events = schedule(stimuli, steps)
    stim_all = (np.unique(np.concatenate([e[1] for e in events]))
                if events else None)

    engine = build_engine(weights=weights, neurons=neurons, silenced=silenced,
                          params=p, device=device, batch=B, seed=config.seed,
                          no_refractory=stim_all, graphs=False,
                          spmv=getattr(config, "spmv", None),
                          road=pick_road(B, getattr(config, "road", "auto")))
    sim = engine._sim
    sim.pop_counts = torch.zeros(n, dtype=torch.int32, device=sim.device)
    sim._sync_struct()
    t_setup = time.perf_counter() - t0
    tape_drive = None
    if tape is not None:
        if isinstance(tape, Tape):
            tape_drive = tape
        elif isinstance(tape, (str, os.PathLike)):
            tape_drive = Tape.load(sim, tape)  # a save_tape file
        else:
            tape_drive = Tape(sim, *tape)

    rates = np.zeros(n, np.float32)  # identical across trials (sim.py:1062)
    counts = np.zeros(n, np.int64)
    binned = np.zeros((n_bins, n), np.int16)
    drive = None
    ev = 0
    change = sorted({e[0] for e in events})
    t1 = time.perf_counter()
    step = 0
    while step < steps:
        while ev < len(events) and events[ev][0] == step:
            _, idx, rate = events[ev]
            rates[idx] = rate
            ev += 1
            if ev == len(events) or events[ev][0] != step:
                drive = _drive_for(sim, rates)
        nxt = min([c for c in change if c > step] + [steps])
        bin_end = (step // bin_steps + 1) * bin_steps
        W = min(L.MAX_WINDOW, nxt - step, bin_end - step, steps - step)
        sim.window(W, tape_drive if tape_drive is not None else drive)
        engine.calls += W
        step += W
        if step % bin_steps == 0 or step == steps:
            b = min((step - 1) // bin_steps, n_bins - 1)
            acc = sim.pop_counts.cpu().numpy().astype(np.int64)
            whole = acc.astype(np.int16)
            binned[b] = whole
            counts += acc
            sim.pop_counts.zero_()
            if on_bin is not None:
                on_bin(b, step * DT, whole)
            if should_stop is not None and should_stop():
                steps = step
                break
    torch.cuda.synchronize()
    t_sim = time.perf_counter() - t1
    cfg_t = steps * DT / 1000.0
    timing = {
        "setup_s": round(t_setup, 3),
        "sim_s": round(t_sim, 3),
        "steps": steps,
        "device": str(device),
        "realtime_ratio": round(cfg_t / t_sim, 4) if t_sim else None,
        "speed": getattr(config, "speed", FAST),
        "graphs": False,
        "windows": "up to %d steps per launch pair" % L.MAX_WINDOW,
        **engine.spmv_info(),
    }
    res = Result(counts, binned, config, list(stimuli), silenced, neurons,
                 timing)
    res.groups = engine._sim.net.conn.groups  # regions of region_timeline
    return res


STATE_FORMAT = "flybrain_amd.engine_state/1"


def save_state(state, path):
    """Write an :meth:`Engine.snapshot` dict as ``.npz`` (arrays only, no
    pickled objects), like the reference's saved resume states
    (lab/flybrain/replay.py:288-300)."""
    arrays = {"format": np.array(STATE_FORMAT)}
    arrays.update({k: np.asarray(v) for k, v in state.items()})
    with open(path, "wb") as f:
        np.savez_compressed(f, **arrays)


def load_state(path):
    """Read a snapshot written by :func:`save_state`; ready for
    :meth:`Engine.restore`."""
    with np.load(path, allow_pickle=False) as z:
        if "format" not in z.files or str(z["format"]) != STATE_FORMAT:
            raise ValueError("%s is not a %s file" % (path, STATE_FORMAT))
        return {k: z[k].copy() for k in z.files if k != "format"}


def _drive_for(sim, rates):
    """Drive table of the neurons with a nonzero rate, all trials."""
    on = np.nonzero(rates > 0)[0]
    if on.size == 0:
        return None
    B = sim.batch
    ne = np.repeat(on, B)
    tr = np.tile(np.arange(B), on.size)
    gr = np.repeat(np.arange(on.size), B)
    return DriveTable(sim, ne, tr, gr, rates[on])


def save(result: Result, out_dir, min_rate=0.5):
    """Write meta.json, rates.parquet and bins.npz into ``out_dir``, the
    reference's files (sim.py:1120-1140): the reference's ``load`` reads
    them back.

    ``rates.parquet`` is :meth:`Result.rate_table` (index = neuron).
    ``bins.npz`` has the reference's keys: ``neuron_idx`` int32 (active
    neurons), ``bins`` int16 [n_bins, n_active], ``bin_ms``, ``t0_ms``.
    """
    out_dir = Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    (out_dir / "meta.json").write_text(
        json.dumps(result.meta(with_indices=True), ensure_ascii=False,
                   indent=2, default=_json_default), encoding="utf-8")
    keep = result.active(min_rate)
    result.rate_table(min_rate).to_parquet(out_dir / "rates.parquet")
    np.savez_compressed(
        out_dir / "bins.npz", neuron_idx=keep.astype(np.int32),
        bins=result.bins[:, keep].astype(np.int16),
        bin_ms=np.float32(result.config.bin_ms),
        t0_ms=np.float32((result.timing or {}).get("запись с, мс", 0.0)))
    return out_dir


def load(out_dir, neurons=None) -> Result:
    """Rebuild a :class:`Result` from a saved run (sim.py:1143-1174)."""
    out_dir = Path(out_dir)
    meta = json.loads((out_dir / "meta.json").read_text(encoding="utf-8"))
    n = meta["n_neurons"] if neurons is None else len(neurons)
    if neurons is None:
        neurons = np.arange(n)
    counts = np.zeros(n, dtype=np.int64)
    if (out_dir / "rates.parquet").exists():
        import pandas as pd

        rates_df = pd.read_parquet(out_dir / "rates.parquet")
        counts[rates_df.index.to_numpy()] = rates_df["n_spikes"].to_numpy()
    else:  # runs saved by an earlier version of this engine
        r = np.load(out_dir / "rates.npz", allow_pickle=False)
        counts[r["neuron_idx"]] = r["n_spikes"]
    z = np.load(out_dir / "bins.npz", allow_pickle=False)
    bins = np.zeros((z["bins"].shape[0], n), dtype=np.int16)
    bins[:, z["neuron_idx"]] = z["bins"]
    cfg = RunConfig(**meta["config"], params=meta["params"])
    stimuli = [Stimulus(indices=np.asarray(s.get("indices", []), np.int64),
                        rate_hz=s["rate_hz"], t_on_ms=s["t_on_ms"],
                        t_off_ms=s["t_off_ms"], name=s["name"])
               for s in meta["stimuli"]]
    silenced = np.asarray(meta.get("silenced") or [], dtype=np.int64)
    return Result(counts, bins, cfg, stimuli, silenced, neurons,
                  meta["timing"])


def _json_default(o):
    if isinstance(o, (np.integer,)):
        return int(o)
    if isinstance(o, (np.floating,)):
        return float(o)
    if isinstance(o, np.ndarray):
        return o.tolist()
    raise TypeError(type(o))
