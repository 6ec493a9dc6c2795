"""The oracle: CPU restatement of the reference LIF step (ctypes over C).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` import this module, and only as the
checker; the engine never does.

``OracleEngine`` mirrors the state and step of ``Engine``
(lab/flybrain/sim.py:348-434, :848-953) on trial-major [B, N] numpy arrays.
The arithmetic lives in ``lif_oracle.c``; this file only derives the
constants the way the reference does (sim.py:361-383) and owns the arrays.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liblif_oracle.so"

MODEL_PARAMS = {  # lab/flybrain/sim.py:36-47
    "tauSyn": 5.0, "tDelay": 1.8, "v0": -52.0, "vReset": -52.0,
    "vRest": -52.0, "vThreshold": -45.0, "tauMem": 20.0, "tRefrac": 2.2,
    "scalePoisson": 250, "wScale": 0.275,
}
DT = 0.1

_lib = None


class _Params(ctypes.Structure):
    _fields_ = [(k, ctypes.c_float) for k in
                ("w_scale", "a_syn", "a_mem", "v_rest", "v_reset", "v_th")]


def build():
    """Compile the oracle with the recipe next to it."""
    import subprocess

    subprocess.run(["make", "-s", "-C", str(HERE), "liblif_oracle.so"],
                   check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        P = ctypes.c_void_p
        L.oracle_philox.argtypes = [P, P, P]
        L.oracle_philox.restype = None
        L.oracle_draw.argtypes = [ctypes.c_uint64, ctypes.c_uint32,
                                  ctypes.c_uint32, ctypes.c_uint32,
                                  ctypes.c_float]
        L.oracle_draw.restype = ctypes.c_int
        L.oracle_stim.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P,
                                  ctypes.c_uint32, ctypes.c_float,
                                  ctypes.c_float, ctypes.c_float, P]
        L.oracle_stim.restype = None
        L.oracle_stim_tape.argtypes = [ctypes.c_int, ctypes.c_int, P,
                                       ctypes.c_float, ctypes.c_float, P]
        L.oracle_stim_tape.restype = None
        L.oracle_step.argtypes = ([ctypes.c_int, ctypes.c_int, P, P, P, P]
                                  + [P] * 5 + [ctypes.c_int, ctypes.c_int, P,
                                               ctypes.c_int] + [P] * 9)
        L.oracle_step.restype = ctypes.c_int
        L.oracle_uniform_grid.argtypes = [ctypes.c_uint64, ctypes.c_uint32,
                                          ctypes.c_uint32, ctypes.c_int,
                                          ctypes.c_uint32, ctypes.c_int, P]
        L.oracle_uniform_grid.restype = None
        L.oracle_run_counts.argtypes = ([ctypes.c_int, ctypes.c_int, P, P, P,
                                         P, P, ctypes.c_float, P, P,
                                         ctypes.c_float, ctypes.c_int, P])
        L.oracle_run_counts.restype = ctypes.c_int
        L.oracle_abi.restype = ctypes.c_int
        if L.oracle_abi() != 2:
            raise RuntimeError("stale %s: rebuild the oracle" % LIB_PATH)
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def philox(ctr, key):
    """Philox4x32-10 of one 4-word counter under a 2-word key."""
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().oracle_philox(_p(c), _p(k), _p(out))
    return out


def draw(seed, trial_id, neuron, t, p):
    return int(lib().oracle_draw(int(seed), int(trial_id), int(neuron),
                                 int(t), float(np.float32(p))))


def uniform_grid(seed, trial_id, n0, n, t0, steps):
    """The drive stream's 24-bit uniforms, uint32 [steps, n] (neurons n0..,
    steps t0..): the values oracle_draw compares with p * 2^24."""
    out = np.zeros((steps, n), np.uint32)
    lib().oracle_uniform_grid(int(seed), int(trial_id), int(n0), int(n),
                              int(t0), int(steps), _p(out))
    return out


def run_counts(conn, steps, tape_ptr, tape_idx, no_refractory=None,
               params=None):
    """Spike counts int64 [n] of one trial run from rest for ``steps``
    steps with the kicks of a sparse tape (neurons tape_idx[tape_ptr[t]:
    tape_ptr[t+1]] at step t): OracleEngine's step, looped in C.  The call
    releases the GIL, so trials run in parallel threads."""
    c = derived(params)
    rs = np.full(conn.n, c["refrac_steps"], np.float32)
    if no_refractory is not None and len(no_refractory):
        rs[np.asarray(no_refractory, dtype=np.int64)] = 0.0
    prm = _Params(*(float(c[k]) for k in ("w_scale", "a_syn", "a_mem",
                                          "v_rest", "v_reset", "v_th")))
    ptr = np.ascontiguousarray(tape_ptr, dtype=np.int64)
    idx = np.ascontiguousarray(tape_idx, dtype=np.int32)
    if ptr.size != steps + 1:
        raise ValueError("tape_ptr needs steps + 1 entries")
    counts = np.zeros(conn.n, np.int64)
    rc = lib().oracle_run_counts(
        conn.n, int(steps), _p(conn.offsets), _p(conn.cols), _p(conn.vals),
        ctypes.byref(prm), _p(rs), float(c["v0"]), _p(ptr),
        _p(idx if idx.size else np.zeros(1, np.int32)), float(c["gain"]),
        c["buf_len"], _p(counts))
    if rc != 0:
        raise ArithmeticError("oracle_run_counts failed (%d)" % rc)
    return counts


def derived(params=None):
    """float32 constants exactly as the reference's tensor ops see them.

    Each is a Python double computed as in sim.py:361-383 and then rounded
    to float32, which is what a float32 tensor op does with a Python scalar.
    """
    p = dict(params or MODEL_PARAMS)
    f = np.float32
    tau_syn_f = DT / p["tauSyn"]
    tau_mem_f = DT / p["tauMem"]
    return {
        "w_scale": f(p["wScale"]),
        "a_syn": f(1.0 - tau_syn_f),
        "a_mem": f(tau_mem_f),
        "v_rest": f(p["vRest"]),
        "v_reset": f(p["vReset"]),
        "v_th": f(p["vThreshold"]),
        "v0": f(p["v0"]),
        "gain": f(p["scalePoisson"]),
        "prob_scale": f(DT / 1000.0),
        "refrac_steps": f(float(round(p["tRefrac"] / DT))),
        "buf_len": int(p["tDelay"] / DT) + 1,
    }


class OracleEngine:
    """State and step of the reference engine, trial-major [B, N]."""

    def __init__(self, conn, batch=1, params=None, no_refractory=None,
                 seeds=None, trial_ids=None, no_refractory_trials=None):
        self.conn = conn
        self.n = n = conn.n
        self.batch = B = int(batch)
        self.c = derived(params)
        self.buf_len = self.c["buf_len"]
        self.refrac_steps = np.full(n, self.c["refrac_steps"], np.float32)
        if no_refractory is not None and len(no_refractory):
            self.refrac_steps[np.asarray(no_refractory, dtype=np.int64)] = 0.0
        if no_refractory_trials is not None:
            # per-trial exemption: B independent single-trial references
            rs = np.tile(self.refrac_steps, (B, 1)).astype(np.float32)
            for b, idx in enumerate(no_refractory_trials):
                if len(idx):
                    rs[b, np.asarray(idx, dtype=np.int64)] = 0.0
            self.refrac_steps = rs
        self.v = np.full((B, n), self.c["v0"], np.float32)
        self.conductance = np.zeros((B, n), np.float32)
        self.spikes = np.zeros((B, n), np.float32)
        self.refrac = (np.tile(self.refrac_steps, (B, 1)).astype(np.float32)
                       if self.refrac_steps.ndim == 1
                       else self.refrac_steps.copy())
        self.delay_buffer = np.zeros((B, self.buf_len, n), np.float32)
        self.ptr = 0
        self.m_fire = self.m_out = self.m_in = self.m_hold = None
        self.v_hold = None
        self.seeds = np.zeros(B, np.uint64) if seeds is None else \
            np.ascontiguousarray(seeds, dtype=np.uint64)
        self.trial_ids = np.arange(B, dtype=np.uint32) if trial_ids is None \
            else np.ascontiguousarray(trial_ids, dtype=np.uint32)
        self.t = 0
        self.rec = np.zeros((B, n), np.float32)  # last synaptic input
        self.icount = np.zeros((B, n), np.int32)
        self._acc = np.zeros(B * n, np.int64)
        self._prm = _Params(*(float(self.c[k]) for k in
                              ("w_scale", "a_syn", "a_mem", "v_rest",
                               "v_reset", "v_th")))

    # masks with the semantics of Engine.set_masks (sim.py:510-571)
    def set_masks(self, fire=None, out=None, inc=None, hold=None,
                  v_clamp=None, trials=None):
        B, n = self.batch, self.n
        rows = slice(None) if trials is None else np.asarray(trials)
        for name, idx in (("m_fire", fire), ("m_out", out), ("m_in", inc),
                          ("m_hold", hold)):
            if idx is None:
                continue
            m = getattr(self, name)
            if m is None:
                m = np.ones((B, n), np.float32)
            m[rows] = 1.0
            if len(idx):
                m[rows, np.asarray(idx, dtype=np.int64)] = 0.0
            setattr(self, name, m)
        if hold is not None:
            clamp = self.c["v_reset"] if v_clamp is None else np.float32(v_clamp)
            if self.v_hold is None:
                self.v_hold = np.zeros((B, n), np.float32)
            self.v_hold[rows] = 0.0
            if len(hold):
                self.v_hold[rows, np.asarray(hold, dtype=np.int64)] = clamp

    def stim_from_rates(self, rates):
        """Poisson drive for the current step from rates [B, N] in Hz."""
        rates = np.ascontiguousarray(rates, dtype=np.float32)
        stim = np.empty((self.batch, self.n), np.float32)
        lib().oracle_stim(self.n, self.batch, _p(rates), _p(self.seeds),
                          _p(self.trial_ids), self.t,
                          float(self.c["prob_scale"]), float(self.c["gain"]),
                          float(self.c["w_scale"]), _p(stim))
        return stim

    def stim_from_bits(self, bits):
        bits = np.ascontiguousarray(bits, dtype=np.uint8)
        stim = np.empty((self.batch, self.n), np.float32)
        lib().oracle_stim_tape(self.n, self.batch, _p(bits),
                               float(self.c["gain"]), float(self.c["w_scale"]),
                               _p(stim))
        return stim

    def step(self, stim):
        """One step given the drive voltage stim [B, N]; returns spikes."""
        stim = np.ascontiguousarray(stim, dtype=np.float32)
        cn = self.conn

        def opt(a):
            return None if a is None else _p(np.ascontiguousarray(a))

        rc = lib().oracle_step(
            self.n, self.batch, _p(cn.offsets), _p(cn.cols), _p(cn.vals),
            ctypes.byref(self._prm), _p(self.v), _p(self.conductance),
            _p(self.spikes), _p(self.refrac), _p(self.delay_buffer),
            self.buf_len, self.ptr, _p(self.refrac_steps),
            int(self.refrac_steps.ndim == 2), opt(self.m_fire),
            opt(self.m_out), opt(self.m_in), opt(self.m_hold),
            opt(self.v_hold), _p(stim), _p(self.rec), _p(self.icount),
            _p(self._acc))
        if rc != 0:
            raise ArithmeticError("integer synaptic input reached 2^24")
        self.ptr = (self.ptr + 1) % self.buf_len
        self.t += 1
        return self.spikes

    def step_rates(self, rates):
        return self.step(self.stim_from_rates(rates))

    def step_bits(self, bits):
        return self.step(self.stim_from_bits(bits))
