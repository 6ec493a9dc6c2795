"""Model constants of the whole-brain LIF model and their float32 forms.

``MODEL_PARAMS`` and ``DT`` are the Shiu et al. defaults of the reference
(lab/flybrain/sim.py:36-48).  :func:`derived` computes the per-step
constants the way ``Engine.__init__`` does (sim.py:361-383): each is a
Python double, and a float32 tensor op rounds it to float32 -- so the
kernels receive exactly those float32 values.
"""

from __future__ import annotations

import numpy as np

MODEL_PARAMS = {
    "tauSyn": 5.0,  # ms, synaptic decay
    "tDelay": 1.8,  # ms, axonal delay
    "v0": -52.0,  # mV, initial voltage
    "vReset": -52.0,  # mV
    "vRest": -52.0,  # mV
    "vThreshold": -45.0,  # mV
    "tauMem": 20.0,  # ms, membrane time constant
    "tRefrac": 2.2,  # ms
    "scalePoisson": 250,  # gain on the Poisson input
    "wScale": 0.275,  # mV per synapse
}
DT = 0.1  # ms

# Largest refractory period (in steps) the one-byte counter represents.
MAX_REFRAC_STEPS = 126


def derived(params=None) -> dict:
    """Per-step constants of a full parameter dict (``params`` replaces
    MODEL_PARAMS whole, as ``Engine.__init__`` takes it, sim.py:350: a key
    it lacks raises KeyError)."""
    p = dict(params or MODEL_PARAMS)
    f = np.float32
    refrac_steps = int(round(p["tRefrac"] / DT))
    if not 0 <= refrac_steps <= MAX_REFRAC_STEPS:
        raise ValueError("tRefrac %.3g ms is %d steps; the engine supports "
                         "0..%d" % (p["tRefrac"], refrac_steps,
                                    MAX_REFRAC_STEPS))
    buf_len = int(p["tDelay"] / DT) + 1
    if buf_len != 19:
        raise ValueError("tDelay %.3g ms gives a %d-slot delay ring; the "
                         "kernels are built for the 19-slot (1.8 ms) ring"
                         % (p["tDelay"], buf_len))
    return {
        "params": p,
        "w_scale": f(p["wScale"]),
        "a_syn": f(1.0 - DT / p["tauSyn"]),
        "a_mem": f(DT / p["tauMem"]),
        "v_rest": f(p["vRest"]),
        "v_reset": f(p["vReset"]),
        "v_th": f(p["vThreshold"]),
        "v0": f(p["v0"]),
        "gain": f(p["scalePoisson"]),
        "prob_scale": f(DT / 1000.0),
        "refrac_steps": refrac_steps,
        "buf_len": buf_len,
    }
