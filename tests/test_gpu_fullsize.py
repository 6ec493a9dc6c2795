"""The full-size synthetic connectome (138,639 neurons, 15,091,983 edges).

* The plain workload (20 neurons at 150 Hz, B = 1) from a recorded Poisson
  tape against the oracle: raster identical over the first 100 ms, v
  bit-exact at 100 ms, per-neuron spike counts over 1 s identical.
* Configs 3 and 4, and the plain drive in 256 trials on the gated road,
  replayed from captured graphs against the oracle, trial by trial, bit for
  bit; the Philox drive's rates against a numpy
  Bernoulli oracle (Pearson >= 0.99); serving on the full connectome.
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


def _oracle_for_trial(conn, wl, b):
    """The oracle running trial ``b`` of a batched workload alone: the
    trial's Philox seed (trial id 0, as the batch runs it), its driven
    neurons exempt from refractoriness, its m_out mask."""
    d = wl.drive_trial == b
    ora = O.OracleEngine(conn, 1, seeds=[wl.seeds[b]], trial_ids=[0],
                         no_refractory=np.unique(wl.drive_neuron[d]))
    o = wl.out_trial == b
    if o.any():
        ora.set_masks(out=list(wl.out_neuron[o]))
    return ora, wl.drive_neuron[d], wl.drive_group[d]


@pytest.mark.parametrize("maker,trials,gate",
                         [(WL.sensory_b256, (0, 97, 255), False),
                          (WL.trials_b64, (3, 63), False),
                          (WL.plain_replicated, (0, 255), True)])
def test_batched_graph_replay_matches_oracle_full_size(full, maker, trials,
                                                       gate):
    """Configs 3 and 4, and the plain drive in 256 trials on the gated road
    (planning launch + work list), at full size as the bench runs them: the
    batch's windows replayed from a captured HIP graph (5 windows per
    replay) for 1,000 steps.  For the named trials every step's raster and
    the final v and g equal, bit for bit, the oracle running that trial
    alone with its seed, rate schedule (10 ms segments for the sensory
    batch), refractory exemptions and m_out mask."""
    conn, net = full
    steps, block = 1000, 100
    wl = maker(conn.n, seed=0, steps=steps)
    run = BatchRunner(net, wl, counts=True, raster_len=block, road="pull",
                      gate=gate)
    assert run.sim.road == "pull" and run.sim.gate == gate
    run.capture(block // run.window)
    oras = {b: _oracle_for_trial(conn, wl, b) for b in trials}
    rates = {b: np.zeros((1, conn.n), np.float32) for b in trials}
    n_seg = wl.rate_hz.shape[0]
    spikes = {b: 0 for b in trials}
    for a in range(0, steps, block):
        run.advance(block, use_graph=True)  # one graph replay
        assert run.sim.t == a + block
        got = {b: run.sim.trial_raster(b, a, a + block) for b in trials}
        for b, (ora, ne, gr) in oras.items():
            for k in range(block):
                t = a + k
                rates[b][0, ne] = wl.rate_hz[
                    min(t // wl.seg_steps, n_seg - 1), gr]
                sp = ora.step_rates(rates[b])[0]
                assert np.array_equal(got[b][k], sp.astype(bool)), \
                    "trial %d: raster differs at step %d" % (b, t)
                spikes[b] += int(sp.sum())
    v = run.sim.view(run.sim.v).cpu().numpy()
    g = run.sim.view(run.sim.g).cpu().numpy()
    if gate:  # most of the network sat calm: the plan did those windows
        assert int(run.sim.calm.sum().item()) > conn.n // 2
    for b, (ora, _, _) in oras.items():
        assert spikes[b] > 0
        np.testing.assert_array_equal(v[b].view(np.uint32),
                                      ora.v[0].view(np.uint32))
        np.testing.assert_array_equal(g[b].view(np.uint32),
                                      ora.conductance[0].view(np.uint32))


def test_philox_rates_match_a_bernoulli_oracle_full_size(full):
    """Statistical parity of the Philox drive: config 2 (20 neurons at
    150 Hz) at full size for 1 s in 64 trials with their own Philox seeds,
    against the oracle driven by independent numpy Bernoulli tapes at the
    same rate (32 trials, run in threads).  Per-neuron rates over the
    neurons that spiked correlate with Pearson r >= 0.99, and every driven
    neuron's count agrees within 5 standard deviations."""
    import concurrent.futures as cf
    import os

    conn, net = full
    steps, B, Bo = 10_000, 64, 32
    wl = WL.plain_replicated(conn.n, seed=0, batch=B, steps=steps)
    run = BatchRunner(net, wl, counts=True)
    run.capture(50)
    run.advance(steps)
    gpu = run.counts().sum(axis=0)
    drv = wl.trial(0).drive_neuron
    p = float(np.float32(150.0) * np.float32(1e-4))

    def one(b):
        rng = np.random.default_rng(10_000 + b)
        kicks = rng.random((steps, drv.size)) < p
        t_, j = np.nonzero(kicks)
        ptr = np.zeros(steps + 1, np.int64)
        np.cumsum(np.bincount(t_, minlength=steps), out=ptr[1:])
        return O.run_counts(conn, steps, ptr, drv[j], no_refractory=drv)

    workers = min(16, len(os.sched_getaffinity(0)))
    with cf.ThreadPoolExecutor(workers) as ex:
        ora = sum(ex.map(one, range(Bo)))
    r_gpu, r_ora = gpu / (B * 1.0), ora / (Bo * 1.0)  # Hz
    act = (gpu > 0) | (ora > 0)
    r = np.corrcoef(r_gpu[act], r_ora[act])[0, 1]
    print("Pearson %.5f over %d neurons; driven mean %.2f vs %.2f Hz"
          % (r, act.sum(), r_gpu[drv].mean(), r_ora[drv].mean()))
    assert r >= 0.99
    # driven neurons: each kick is a spike, so counts ~ Binomial(steps, p)
    # per trial; compare rates with the two samples' standard errors
    se = np.sqrt(p * (1 - p) * steps * (1.0 / B + 1.0 / Bo)) / (steps * 1e-4)
    assert np.all(np.abs(r_gpu[drv] - r_ora[drv]) < 5 * se + 1e-9)


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


def test_serving_full_size_alone_and_in_the_mix(full):
    """Continuous batching on the full connectome: 16 requests of 500-2,000
    steps through 8 slots (slots reused, shards of 4 on the push road);
    every result equals the same request served alone, and the request
    admitted last (its slot restarted at a later clock) equals the oracle
    running it from step 0 with the same Philox stream."""
    from flybrain_amd.serve import Server, SimRequest, poisson_requests

    conn, net = full
    reqs = poisson_requests(conn.n, 16, 200.0, seed=9, steps=(500, 2000))
    mix, st = Server(net, slots=8, shard=4, graph_windows=5).serve(
        [SimRequest(**vars(r)) for r in reqs], realtime=False)
    assert st["requests"] == 16
    got = {r.rid: r for r in mix}
    assert len({r.slot for r in mix}) == 8
    assert sum(int(r.counts.sum()) for r in mix) > 0
    for r in reqs:
        alone, _ = Server(net, slots=1, road="push").serve(
            [SimRequest(**vars(r))], realtime=False)
        np.testing.assert_array_equal(got[r.rid].counts, alone[0].counts,
                                      err_msg="request %d" % r.rid)
    late = max(mix, key=lambda x: (x.admitted_s, x.rid))
    r = next(q for q in reqs if q.rid == late.rid)
    ora = O.OracleEngine(conn, 1, seeds=[r.seed], trial_ids=[0],
                         no_refractory=np.unique(r.drive))
    if len(r.silenced):
        ora.set_masks(out=list(r.silenced))
    rates = np.zeros((1, conn.n), np.float32)
    rates[0, r.drive] = r.rate_hz
    counts = np.zeros(conn.n, np.int64)
    for _ in range(r.steps):
        counts += ora.step_rates(rates)[0].astype(np.int64)
    np.testing.assert_array_equal(late.counts, counts)
    assert counts.sum() > 0
