"""The full-size synthetic connectome (138,639 neurons, 15,091,983 edges).

* The plain workload (20 neurons at 150 Hz, B = 1) from a recorded Poisson
  tape against the oracle: raster identical over the first 100 ms, v
  bit-exact at 100 ms, per-neuron spike counts over 1 s identical (Pearson
  1 >= 0.99).
* Batched trials: trials of the B = 256 sensory batch (pull road) and of
  the B = 64 trial batch give exactly the raster of the same trial run
  alone (B = 1, push road).
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from flybrain_amd import connectome as C
from flybrain_amd import workloads as WL
from flybrain_amd.batch import BatchRunner
from flybrain_amd.device import DeviceNet, DeviceSim, Tape
from gpu_helpers import state_of

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def full():
    conn = C.full()
    return conn, DeviceNet(conn)


def test_plain_b1_full_size_against_oracle(full):
    conn, net = full
    wl = WL.plain_b1(conn.n, seed=0)
    drv = wl.drive_neuron
    steps, check = 10_000, 1_000
    rng = np.random.default_rng(11)
    p = float(np.float32(150.0) * np.float32(1e-4))
    kicks = rng.random((steps, drv.size)) < p
    t_, k_ = np.nonzero(kicks)
    sim = DeviceSim(net, 1, no_refractory=drv, raster_len=check,
                    counts=True)
    tape = Tape(sim, t_, drv[k_])
    ora = O.OracleEngine(conn, 1, no_refractory=drv)
    bits = np.zeros((1, conn.n), np.uint8)
    for a in range(0, check, 20):
        sim.window(20, tape)
    raster = sim.read_raster(0, check)
    counts = np.zeros(conn.n, np.int64)
    for t in range(steps):
        bits[:] = 0
        bits[0, drv[kicks[t]]] = 1
        sp = ora.step_bits(bits)
        counts += sp[0].astype(np.int64)
        if t < check:
            assert np.array_equal(raster[t, 0], sp[0].astype(bool)), t
        if t == check - 1:
            v, g, _, _ = state_of(sim)
            np.testing.assert_array_equal(v.view(np.uint32),
                                          ora.v.view(np.uint32))
            np.testing.assert_array_equal(g.view(np.uint32),
                                          ora.conductance.view(np.uint32))
    for a in range(check, steps, 20):
        sim.window(20, tape)
    got = sim.view(sim.counts).cpu().numpy()[0].astype(np.int64)
    np.testing.assert_array_equal(got, counts)
    assert counts.sum() > 0
    act = counts > 0
    r = np.corrcoef(got[act], counts[act])[0, 1] if act.sum() > 1 else 1.0
    assert r >= 0.99


@pytest.mark.parametrize("maker,trials", [(WL.sensory_b256, (0, 97, 255)),
                                          (WL.trials_b64, (3, 63))])
def test_batched_trial_equals_trial_alone(full, maker, trials):
    conn, net = full
    steps = 300
    wl = maker(conn.n, seed=0, steps=steps)
    batch = BatchRunner(net, wl, counts=True, raster_len=steps)
    assert batch.sim.road == "pull"
    batch.advance(steps, use_graph=False)
    got = batch.sim.read_raster(0, steps)
    assert got.any()
    for b in trials:
        one = BatchRunner(net, wl.trial(b), counts=True, raster_len=steps,
                          road="push")
        one.advance(steps, use_graph=False)
        want = one.sim.read_raster(0, steps)[:, 0]
        assert np.array_equal(got[:, b], want), "trial %d" % b
        assert want.any()
