"""fis_window_noalloc: the push-road window (fis_update + fis_scatter) with
its event staging in the library's static device scratch instead of the
caller's buffers -- for callers that cannot allocate.  Checked against
fis_update + fis_scatter on identical simulations, window by window, on
two simulations in a row (the second smaller, so it runs over what the
first left in the scratch), and its size limit on the host."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from flybrain_amd import _lib as L
from flybrain_amd import connectome as C
from flybrain_amd.device import DeviceNet, DeviceSim, DriveTable


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _make(net, B, seeds, drv, rate):
    sim = DeviceSim(net, B, seeds=seeds, raster_len=40, counts=True,
                    road="push")
    ne = np.concatenate(drv)
    tr = np.concatenate([np.full(len(d), b) for b, d in enumerate(drv)])
    sim.exempt_pairs(ne, tr)
    return sim, DriveTable(sim, ne, tr, tr, np.array(rate, np.float32))


def _compare(a, b, where):
    for name in ("v", "g", "st", "ring", "counts", "raster"):
        x, y = getattr(a, name), getattr(b, name)
        assert torch.equal(x, y), "%s: %s differs" % (where, name)


@pytest.mark.gpu
def test_noalloc_window_equals_update_then_scatter():
    cases = [  # (neurons, trials, windows): the second input is smaller
        (C.synthetic(3000, None, mean_out=60, count_exp=1.2, n_groups=5,
                     seed=41), 64, [20, 20, 7, 20, 1, 20, 20, 13]),
        (C.synthetic(700, None, mean_out=40, count_exp=1.2, n_groups=3,
                     seed=42), 3, [20, 5, 20, 20, 1, 20]),
    ]
    for conn, B, windows in cases:
        net = DeviceNet(conn)
        rng = np.random.default_rng(B)
        seeds = rng.integers(0, 2**63, size=B, dtype=np.uint64)
        drv = [rng.choice(conn.n, int(rng.integers(5, 60)), replace=False)
               for _ in range(B)]
        rate = [float(rng.uniform(100, 900)) for _ in range(B)]
        a, ta = _make(net, B, seeds, drv, rate)
        b, tb = _make(net, B, seeds, drv, rate)
        spikes = 0
        for k, W in enumerate(windows):
            a.drive_table(ta, W)
            a.update(W)
            a.scatter()
            b.drive_table(tb, W)
            L.check(b.lib.fis_window_noalloc(b._sp, W, _stream()),
                    "fis_window_noalloc")
            b._moved(W)
            torch.cuda.synchronize()
            _compare(a, b, "n=%d B=%d window %d" % (conn.n, B, k))
            spikes = int(a.counts.sum())
        assert spikes > 0
        assert int(a.clock.item()) == int(b.clock.item()) == sum(windows)


def test_noalloc_rejects_what_its_scratch_cannot_hold():
    """The size check runs on the host, before any device call: a
    simulation whose staging exceeds FIS_SCRATCH_BYTES is refused."""
    lib = L.lib()
    s = L.FisSim()
    s.n, s.batch, s.bstride = 2_000_000, 64, 64  # ~13 M neuron-trials
    for name in ("offsets", "cols", "vals", "v", "g", "st", "fl", "ring",
                 "sbits", "clock", "seeds", "trial_ids"):
        setattr(s, name, 64)  # never dereferenced: the call is refused
    s.refrac_steps = 22
    invalid = 1  # hipErrorInvalidValue, not a device error: no device call
    assert lib.fis_window_noalloc(ctypes.byref(s), 20, None) == invalid
    assert lib.fis_window_noalloc(ctypes.byref(s), 0, None) == invalid
    s.n, s.batch, s.bstride = 1_000_000, 1, 1    # 322 B per neuron-trial
    assert lib.fis_window_noalloc(ctypes.byref(s), 20, None) == invalid
