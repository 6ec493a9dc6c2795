"""Shared helpers of the GPU parity tests: run the HIP path and the oracle
side by side on the same inputs and compare what the reference defines."""

from __future__ import annotations

import numpy as np

from flybrain_amd import connectome as C


def tiny_conn(z):
    return C.Connectome(int(z["n"]), z["conn_offsets"], z["conn_cols"],
                        z["conn_vals"])


def mask_schedule(z):
    """{step: set_masks kwargs} of the tiny_masks fixture."""
    out = {}
    for k, s in enumerate(z["sched_steps"]):
        kw = {}
        for name in ("fire", "out", "inc", "hold", "v_clamp"):
            key = "sched%d_%s" % (k, name)
            if key in z.files:
                kw[name] = float(z[key]) if name == "v_clamp" else z[key]
        out[int(s)] = kw
    return out


def apply_masks_device(sim, kw, v_reset=-52.0):
    """Engine.set_masks semantics (sim.py:510-571) on a DeviceSim."""
    from flybrain_amd import _lib as L

    for bit, key in ((L.FIRE_OFF, "fire"), (L.OUT_OFF, "out"),
                     (L.IN_OFF, "inc")):
        if key in kw and kw[key] is not None:
            sim.set_flag(bit, kw[key])
    if "hold" in kw and kw["hold"] is not None:
        clamp = kw.get("v_clamp")
        sim.set_hold(kw["hold"], v_reset if clamp is None else clamp)


def state_of(sim):
    """(v, g, spikes, refrac counter) as numpy [B, n] from the device."""
    v = sim.view(sim.v).cpu().numpy()
    g = sim.view(sim.g).cpu().numpy()
    sp = sim.spike_state().cpu().numpy().astype(np.float32)
    r = sim.refrac_counter().cpu().numpy().astype(np.float32)
    return v, g, sp, r


def compare_state(sim, ora, where):
    v, g, sp, r = state_of(sim)
    assert np.array_equal(sp, ora.spikes), "%s: spikes differ" % where
    np.testing.assert_array_equal(v.view(np.uint32), ora.v.view(np.uint32),
                                  err_msg="%s: v not bit-exact" % where)
    np.testing.assert_array_equal(g.view(np.uint32),
                                  ora.conductance.view(np.uint32),
                                  err_msg="%s: g not bit-exact" % where)
    want_r = np.minimum(ora.refrac, 127).astype(np.float32)
    np.testing.assert_array_equal(r, want_r,
                                  err_msg="%s: refractory counter" % where)
