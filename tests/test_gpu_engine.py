"""The drop-in engine surface on the GPU: Engine.step / use_graphs / masks /
snapshot-restore / run, against the oracle and the reference's golden
vectors, through the C ABI."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
from flybrain_amd import connectome as C
from flybrain_amd import sim as S
from flybrain_amd.device import DeviceNet
from gpu_helpers import tiny_conn

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("road", ["push", "pull"])
def test_engine_step_matches_oracle_philox(road):
    """Engine.step(rates) with masks changing, B=3, against the oracle fed
    the same Philox stream; the graph road gives the same bytes."""
    conn = C.synthetic(400, None, mean_out=40, count_exp=1.2, n_groups=3,
                       seed=8)
    n, B = conn.n, 3
    drv = np.arange(0, n, 9)
    rates = torch.zeros(B, n, device="cuda")
    rates[:, drv] = 600.0
    rates[1, drv[::2]] = 2500.0
    outs = []
    for graphs in (False, True):
        eng = S.Engine(conn, n, batch=B, seed=1234, no_refractory=drv,
                       road=road)
        if graphs:
            assert not eng.use_graphs()  # declined: direct steps are faster
            assert eng.use_graphs(force=True)
        ora = O.OracleEngine(conn, B, no_refractory=drv,
                             seeds=[1234] * B)
        for t in range(300):
            if t == 100:
                eng.set_masks(fire=[5, 6], out=[10, 11, 12], inc=[20],
                              hold=[30, 31], v_clamp=-58.0)
                ora.set_masks(fire=[5, 6], out=[10, 11, 12], inc=[20],
                              hold=[30, 31], v_clamp=-58.0)
            if t == 200:
                eng.set_masks(fire=[], out=[], inc=[], hold=[])
                ora.set_masks(fire=[], out=[], inc=[], hold=[])
            sp = eng.step(rates)
            ora.step_rates(rates.cpu().numpy())
            assert np.array_equal(sp.cpu().numpy(), ora.spikes), "t=%d" % t
        np.testing.assert_array_equal(eng.v.cpu().numpy().view(np.uint32),
                                      ora.v.view(np.uint32))
        np.testing.assert_array_equal(
            eng.conductance.cpu().numpy().view(np.uint32),
            ora.conductance.view(np.uint32))
        db = eng.delay_buffer.cpu().numpy()
        np.testing.assert_array_equal(db + np.float32(0), ora.delay_buffer
                                      + np.float32(0))
        assert eng.ptr == ora.ptr and eng.calls == 300
        outs.append(eng.v.cpu().numpy().copy())
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("road", ["push", "pull"])
def test_snapshot_restore_continues_bit_exact(road):
    """Resume from a snapshot into a fresh engine: the continuation is the
    same bytes (probe/resume_state.py); dropping the ring breaks it."""
    conn = C.synthetic(500, None, mean_out=40, count_exp=1.2, n_groups=3,
                       seed=6)
    n, B = conn.n, 2
    drv = np.arange(0, n, 11)
    rates = torch.zeros(B, n, device="cuda")
    rates[:, drv] = 900.0
    a = S.Engine(conn, n, batch=B, seed=77, no_refractory=drv, road=road)
    for _ in range(137):
        a.step(rates)
    snap = a.snapshot()
    ref = []
    for _ in range(120):
        ref.append(a.step(rates).cpu().numpy().copy())
    b = S.Engine(conn, n, batch=B, seed=5, no_refractory=drv, road=road)
    done = b.restore(snap)
    assert {"v", "conductance", "spikes", "refrac", "delay_buffer", "ptr",
            "gen"} <= set(done)
    for k in range(120):
        assert np.array_equal(b.step(rates).cpu().numpy(), ref[k]), k
    c = S.Engine(conn, n, batch=B, seed=5, no_refractory=drv, road=road)
    c.restore(snap, drop=("delay_buffer",))
    differs = any(not np.array_equal(c.step(rates).cpu().numpy(), ref[k])
                  for k in range(120))
    assert differs, "dropping the ring must change the continuation"


@pytest.mark.parametrize("road", ["push", "pull"])
def test_run_matches_reference_golden(golden, road):
    """run() fed the reference's recorded Bernoulli tape: per-bin int16
    raster and per-neuron counts equal the reference's run()."""
    z = golden("run_sched.npz")
    conn = C.Connectome(int(z["n"]), z["conn_offsets"], z["conn_cols"],
                        z["conn_vals"])
    n, B = conn.n, int(z["batch"])
    st = [S.Stimulus(indices=z["stim_a"], rate_hz=400.0, t_on_ms=5.0,
                     t_off_ms=32.5, name="a"),
          S.Stimulus(indices=z["stim_b"], rate_hz=900.0, t_on_ms=20.0,
                     t_off_ms=None, name="b")]
    tape = z["tape_bits"]  # [steps, B, packed n]
    bits = np.unpackbits(tape, axis=-1)[..., :n]
    t, b, i = np.nonzero(bits)
    cfg = S.RunConfig(t_run_s=float(z["t_run_s"]), bin_ms=float(z["bin_ms"]),
                      n_trials=B, seed=0, device="cuda", road=road)
    got_bins = []
    res = S.run(st, silenced=z["silenced"], config=cfg, weights=conn,
                tape=(t, i, b), on_bin=lambda k, ms, c: got_bins.append(k))
    np.testing.assert_array_equal(res.counts, z["counts"])
    np.testing.assert_array_equal(res.bins, z["bins"])
    assert got_bins == list(range(z["bins"].shape[0]))


def test_engine_refuses_cpu_and_bad_roads():
    conn = C.synthetic(50, None, mean_out=5, n_groups=0, seed=1)
    with pytest.raises(S.KernelUnavailable):
        S.Engine(conn, 50, device="cpu")
    with pytest.raises(ValueError):
        S.Engine(conn, 50, spmv="nope")


@pytest.mark.parametrize("road", ["push", "pull"])
def test_tiny_masks_through_engine(golden, road):
    """The Engine surface on the tiny fixture (step_tape, set_masks): v,
    g and spikes fingerprints equal the reference's at every step."""
    import hashlib

    from flybrain_amd.device import Tape
    from gpu_helpers import mask_schedule

    z = golden("tiny_masks.npz")
    conn = tiny_conn(z)
    n, B = conn.n, int(z["batch"])
    bits = np.unpackbits(z["tape_bits"], axis=-1)[..., :n]
    eng = S.Engine(conn, n, batch=B, no_refractory=z["no_refractory"],
                   road=road)
    t_, b_, i_ = np.nonzero(bits)
    tape = Tape(eng._sim, t_, i_, b_)
    sched = mask_schedule(z)

    def fp(a):
        return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]
    for t in range(int(z["steps"])):
        if t in sched:
            eng.set_masks(**sched[t])
        sp = eng.step_tape(tape)
        assert fp(eng.v.cpu().numpy()) == z["fps"][t][0], t
        assert fp(eng.conductance.cpu().numpy()) == z["fps"][t][1], t
        assert fp(sp.cpu().numpy()) == z["fps"][t][2], t
        db = eng.delay_buffer.cpu().numpy() + np.float32(0)
        assert fp(db) == z["fps"][t][4], t


@pytest.mark.parametrize("road", ["push", "pull"])
def test_out_mask_masks_the_previous_steps_spikes(road):
    """m_out set (or lifted) between steps applies to the spikes of the step
    just taken, as rec = W @ (spikes * m_out) in the reference's next step
    (sim.py:910-918): a spike at step T-1 followed by set_masks(out=...)
    before step T is not transmitted (and is, when the mask is lifted)."""
    from flybrain_amd.device import Tape

    conn = C.synthetic(300, None, mean_out=60, count_exp=1.1, n_groups=0,
                       seed=21)
    n, B = conn.n, 2
    hub = int(np.argmax(conn.degree()))
    kicks = {9: [(hub, 0), (hub, 1)], 39: [(hub, 0), (hub, 1)],
             69: [(hub, 1)]}
    steps = 100
    bits = np.zeros((steps, B, n), np.uint8)
    for t, pairs in kicks.items():
        for i, b in pairs:
            bits[t, b, i] = 1
    t_, b_, i_ = np.nonzero(bits)
    eng = S.Engine(conn, n, batch=B, road=road)
    tape = Tape(eng._sim, t_, i_, b_)
    ora = O.OracleEngine(conn, B)
    masks = {10: dict(out=[hub]), 40: dict(out=[]), 70: dict(out=[hub])}
    for t in range(steps):
        if t in masks:
            eng.set_masks(**masks[t])
            ora.set_masks(**masks[t])
        sp = eng.step_tape(tape)
        ora.step_bits(bits[t])
        assert np.array_equal(sp.cpu().numpy(), ora.spikes), t
        np.testing.assert_array_equal(
            eng.conductance.cpu().numpy().view(np.uint32),
            ora.conductance.view(np.uint32), err_msg="t=%d" % t)
        np.testing.assert_array_equal(eng.v.cpu().numpy().view(np.uint32),
                                      ora.v.view(np.uint32),
                                      err_msg="t=%d" % t)
    # the hub did spike at each kicked step (the test exercised the fix)
    assert ora.conductance.any()


def test_roads_agree_after_restore_and_out_mask():
    """Restore a snapshot into a push-road and a pull-road engine, switch
    m_out on spiking hubs at once (their last spikes sit in the restored
    ring), then step: both roads give the same spikes and state."""
    conn = C.synthetic(500, None, mean_out=50, count_exp=1.15, n_groups=3,
                       seed=41)
    n, B = conn.n, 3
    drv = np.arange(0, n, 7)
    rates = torch.zeros(B, n, device="cuda")
    rates[:, drv] = 1500.0
    src = S.Engine(conn, n, batch=B, seed=9, no_refractory=drv, road="push")
    for _ in range(83):
        src.step(rates)
    snap = src.snapshot()
    hubs = list(np.argsort(-conn.degree())[:30])
    engines = [S.Engine(conn, n, batch=B, seed=1, no_refractory=drv,
                        road=road) for road in ("push", "pull")]
    for eng in engines:
        eng.restore(snap)
        eng.set_masks(out=hubs)
        eng.use_graphs(force=True)
    assert engines[1].road == "pull"
    for t in range(60):
        a, b = (e.step(rates).cpu().numpy().copy() for e in engines)
        assert np.array_equal(a, b), t
    np.testing.assert_array_equal(engines[0].v.cpu().numpy().view(np.uint32),
                                  engines[1].v.cpu().numpy().view(np.uint32))
    np.testing.assert_array_equal(
        engines[0].delay_buffer.numpy(), engines[1].delay_buffer.numpy())


def test_tuned_runner_matches_default(tmp_path, monkeypatch):
    """The autotuner times every (road, workgroup shape) candidate on the
    workload itself and the runner built on its choice gives the same
    counts as the default one: the shape changes speed, not results."""
    from flybrain_amd import autotune, workloads as WL
    from flybrain_amd.batch import BatchRunner

    monkeypatch.setenv("FLYBRAIN_TUNE_CACHE", str(tmp_path / "tune.json"))
    conn = C.synthetic(2048, None, mean_out=60, count_exp=1.2, n_groups=6,
                       seed=31)
    net = DeviceNet(conn)
    wl = WL.trials_b64(conn.n, seed=3, batch=96, steps=400)
    res = autotune.tune(net, wl, warmup_steps=40, windows=3)
    assert len(res["times"]) == len(autotune.candidates(96))
    assert all(ms > 0 for ms in res["times"].values())
    assert (tmp_path / "tune.json").exists()
    tuned = BatchRunner(net, wl, counts=True, road="tuned")
    assert tuned.sim.road == res["best"]["road"]
    base = BatchRunner(net, wl, counts=True)
    tuned.advance(400, use_graph=False)
    base.advance(400, use_graph=False)
    np.testing.assert_array_equal(tuned.counts(), base.counts())


@pytest.mark.parametrize("batch,road,fused,gate,sparse",
                         [(1, "push", True, False, False),
                          (3, "push", False, False, False),
                          (64, "push", False, False, False),
                          (64, "pull", False, False, False),
                          (256, "pull", False, False, False),
                          (64, "pull", False, True, False),
                          (256, "pull", False, True, True)])
def test_graph_replays_equal_eager_windows(batch, road, fused, gate, sparse):
    """A captured graph of several windows leaves the device clock alone
    until its one closing tick (each window reads the clock plus the steps
    launched before it).  Replay after replay, with eager windows between,
    it gives the counts, state and clock of the same windows run one by
    one; the drive table's rates change every 7 steps inside a replay.
    Gated, each window is a planning launch and a walk of its work list,
    whose counters alternate from window to window inside the graph and
    out of it; the sparse drive leaves most tiles calm."""
    from flybrain_amd import workloads as WL
    from flybrain_amd.batch import BatchRunner

    # (sparse: few out-edges, so most neurons are never reached and settle)
    conn = C.synthetic(1024, None, mean_out=6 if sparse else 50,
                       count_exp=1.2, n_groups=4, seed=37)
    net = DeviceNet(conn)
    if batch == 1:
        wl = WL.plain_b1(conn.n, seed=5, steps=1000, n_driven=60,
                         rate_hz=400.0)
    elif sparse:
        wl = WL.plain_replicated(conn.n, seed=5, batch=batch, steps=1000,
                                 n_driven=4, rate_hz=300.0)
    else:
        wl = WL.trials_b64(conn.n, seed=5, batch=batch, steps=1000)
    if not fused:  # rates that change every 7 steps, also inside windows
        n_groups = wl.rate_hz.shape[1]
        rng = np.random.default_rng(batch)
        wl.rate_hz = rng.uniform(0.0, 600.0, size=(150, n_groups)).astype(
            np.float32)
        wl.seg_steps = 7
    graph = BatchRunner(net, wl, counts=True, road=road, fused_drive=fused,
                        gate=gate)
    eager = BatchRunner(net, wl, counts=True, road=road, fused_drive=fused,
                        gate=gate)
    assert graph.fused == eager.fused == fused
    assert graph.sim.gate == eager.sim.gate == gate
    graph.capture(3)
    for rep in range(3):
        graph.advance(60)  # one replay of three 20-step windows
        eager.advance(60, use_graph=False)
        torch.cuda.synchronize()
        assert graph.t == eager.t
        assert int(graph.sim.clock.item()) == int(eager.sim.clock.item()) \
            == eager.t, "clock after replay %d" % rep
        np.testing.assert_array_equal(graph.counts(), eager.counts(),
                                      err_msg="counts after replay %d" % rep)
        torch.testing.assert_close(graph.sim.v, eager.sim.v, rtol=0, atol=0)
        graph.advance(20, use_graph=False)  # an eager window between
        eager.advance(20, use_graph=False)
    assert eager.counts().sum() > 0
    if sparse:  # tiles went calm: the planning launch did their windows
        assert int(graph.sim.calm.sum().item()) > 0


@pytest.mark.parametrize("road", ["push", "pull"])
def test_engine_switches_between_graph_and_direct_steps(road):
    """Direct steps run as one deferred-clock run (each launch reads the
    clock plus the steps since it last moved; the clock catches up when a
    graph is captured or replayed): switching between direct and graph
    steps several times keeps every step equal to the oracle's, and the
    device clock ends equal to the step count."""
    conn = C.synthetic(400, None, mean_out=40, count_exp=1.2, n_groups=3,
                       seed=8)
    n, B = conn.n, 3
    drv = np.arange(0, n, 9)
    rates = torch.zeros(B, n, device="cuda")
    rates[:, drv] = 1500.0
    rates[1, drv[::2]] = 3000.0
    eng = S.Engine(conn, n, batch=B, seed=4321, no_refractory=drv, road=road)
    ora = O.OracleEngine(conn, B, no_refractory=drv, seeds=[4321] * B)
    r_np = rates.cpu().numpy()
    t = 0
    for graphs, steps in ((False, 45), (True, 30), (False, 25), (True, 20),
                          (False, 13)):
        assert eng.use_graphs(graphs, force=True) == graphs
        for _ in range(steps):
            sp = eng.step(rates)
            ora.step_rates(r_np)
            assert np.array_equal(sp.cpu().numpy(), ora.spikes), \
                "graphs=%s, step %d" % (graphs, t)
            t += 1
    eng.use_graphs(True, force=True)  # closes the last direct run
    torch.cuda.synchronize()
    assert int(eng._sim.clock.item()) == t == eng._sim.t
    assert ora.spikes.sum() > 0


@pytest.mark.parametrize("road", ["push", "pull"])
def test_rate_cap_and_zero_rates_match_oracle(road):
    """Rates at the cap (10,000 Hz: a kick every step, p = 1.0f) and
    exactly zero, side by side in one batch: every step equals the
    oracle's, and a capped neuron (refractory-exempt) fires every step."""
    conn = C.synthetic(300, None, mean_out=30, count_exp=1.2, n_groups=2,
                       seed=17)
    n, B = conn.n, 2
    drv = np.arange(0, n, 13)
    rates = torch.zeros(B, n, device="cuda")
    rates[0, drv] = 10_000.0  # trial 1 stays undriven
    eng = S.Engine(conn, n, batch=B, seed=99, no_refractory=drv, road=road)
    ora = O.OracleEngine(conn, B, no_refractory=drv, seeds=[99] * B)
    r_np = rates.cpu().numpy()
    for t in range(80):
        sp = eng.step(rates).cpu().numpy()
        ora.step_rates(r_np)
        assert np.array_equal(sp, ora.spikes), "step %d" % t
        assert sp[0, drv].all(), "a capped neuron missed step %d" % t


@pytest.mark.parametrize("road", ["push", "pull"])
def test_network_without_edges_runs(road):
    """Every neuron silenced, or no synapse at all: the engine runs with
    zero recurrent input, as the reference's empty W does; driven neurons
    follow their kicks exactly as the oracle's."""
    conn = C.from_edges(120, [], [], [])
    n, B = conn.n, 2
    drv = np.arange(0, n, 7)
    rates = torch.zeros(B, n, device="cuda")
    rates[:, drv] = 2000.0
    eng = S.Engine(conn, n, batch=B, seed=3, no_refractory=drv, road=road)
    ora = O.OracleEngine(conn, B, no_refractory=drv, seeds=[3] * B)
    r_np = rates.cpu().numpy()
    for t in range(60):
        sp = eng.step(rates).cpu().numpy()
        ora.step_rates(r_np)
        assert np.array_equal(sp, ora.spikes), "step %d" % t
    assert ora.spikes.sum() >= 0 and sp[:, np.setdiff1d(np.arange(n), drv)].sum() == 0


@pytest.mark.parametrize("road,batch", [("push", 3), ("pull", 64)])
def test_recorded_philox_tape_replays_exactly(tmp_path, road, batch):
    """The tape as a file: the kicks a Philox drive table made in a run
    are recorded window by window, written as .npz, read back and fed to
    a fresh simulation through Tape -- the raster and the final state are
    those of the Philox run."""
    from flybrain_amd.device import DeviceNet, DeviceSim, DriveTable, Tape
    from flybrain_amd.device import save_tape

    conn = C.synthetic(800, None, mean_out=50, count_exp=1.2, n_groups=4,
                       seed=8)
    net = DeviceNet(conn)
    rng = np.random.default_rng(batch)
    seeds = rng.integers(0, 2**63, size=batch, dtype=np.uint64)
    ne = np.concatenate([rng.choice(conn.n, 30, replace=False)
                         for _ in range(batch)])
    tr = np.repeat(np.arange(batch), 30)

    def make():
        s = DeviceSim(net, batch, seeds=seeds, raster_len=200, road=road)
        s.exempt_pairs(ne, tr)
        return s

    a = make()
    table = DriveTable(a, ne, tr, tr, rng.uniform(200, 800, batch))
    kicks = [a.record_window(20, table) for _ in range(10)]
    st, nn, tt = (np.concatenate(x) for x in zip(*kicks))
    assert st.size > 0 and st.max() < 200
    path = tmp_path / "tape.npz"
    save_tape(path, st, nn, tt, n=conn.n, batch=batch, seeds=seeds)
    b = make()
    tape = Tape.load(b, path)
    for _ in range(10):
        b.window(20, tape)
    assert np.array_equal(a.read_raster(0, 200), b.read_raster(0, 200))
    assert a.read_raster(0, 200).sum() > st.size  # recurrent spikes too
    for name in ("v", "g", "st"):
        assert torch.equal(getattr(a, name), getattr(b, name)), name


def test_run_takes_a_tape_file(golden, tmp_path):
    """run(tape=<path>) with the reference's recorded tape written as a
    tape file gives the reference's run() result."""
    from flybrain_amd.device import save_tape

    z = golden("run_sched.npz")
    conn = C.Connectome(int(z["n"]), z["conn_offsets"], z["conn_cols"],
                        z["conn_vals"])
    n, B = conn.n, int(z["batch"])
    st = [S.Stimulus(indices=z["stim_a"], rate_hz=400.0, t_on_ms=5.0,
                     t_off_ms=32.5, name="a"),
          S.Stimulus(indices=z["stim_b"], rate_hz=900.0, t_on_ms=20.0,
                     t_off_ms=None, name="b")]
    bits = np.unpackbits(z["tape_bits"], axis=-1)[..., :n]
    t, b, i = np.nonzero(bits)
    save_tape(tmp_path / "tape.npz", t, i, b, n=n, batch=B)
    cfg = S.RunConfig(t_run_s=float(z["t_run_s"]), bin_ms=float(z["bin_ms"]),
                      n_trials=B, seed=0, device="cuda")
    res = S.run(st, silenced=z["silenced"], config=cfg, weights=conn,
                tape=str(tmp_path / "tape.npz"))
    np.testing.assert_array_equal(res.counts, z["counts"])
    np.testing.assert_array_equal(res.bins, z["bins"])


@pytest.mark.parametrize("src_road,dst_road", [("push", "push"),
                                               ("pull", "push"),
                                               ("push", "pull")])
def test_trial_state_file_resumes_in_another_slot(tmp_path, src_road,
                                                  dst_road):
    """A running trial saved (trial_state -> file) at a window boundary and
    loaded into another slot of another simulation, later in that
    simulation's clock and on either road, continues bit for bit as the
    uninterrupted trial: raster of every later step and final state."""
    from flybrain_amd.device import (DeviceNet, DeviceSim, DriveTable,
                                     load_trial_state, save_trial_state)

    conn = C.synthetic(900, None, mean_out=50, count_exp=1.2, n_groups=4,
                       seed=12)
    net = DeviceNet(conn)
    rng = np.random.default_rng(5)
    B = 64 if "pull" in (src_road, dst_road) else 4
    seeds = rng.integers(0, 2**63, size=B, dtype=np.uint64)
    drv = rng.choice(conn.n, 40, replace=False)
    sil = rng.choice(conn.n, 6, replace=False)

    def sim(road, seeds):
        s = DeviceSim(net, B, seeds=seeds, raster_len=200, counts=True,
                      road=road, trial_ids=np.zeros(B, np.uint32))
        s.set_trial_end(np.arange(B), np.full(B, 10**9))
        return s

    def table(s, trials):
        ne = np.tile(drv, len(trials))
        tr = np.repeat(trials, drv.size)
        return DriveTable(s, ne, tr, np.zeros(ne.size, np.int64),
                          np.array([400.0], np.float32))

    src = sim(src_road, seeds)
    src.exempt_pairs(np.tile(drv, B), np.repeat(np.arange(B), drv.size))
    src.set_flag_pairs(2, np.tile(sil, B), np.repeat(np.arange(B), sil.size))
    ts = table(src, np.arange(B))
    for _ in range(6):
        src.window(20, ts)
    k = 2  # the trial moved
    save_trial_state(tmp_path / "trial.npz", src.trial_state(k))
    ref_raster = []
    for _ in range(8):
        src.window(20, ts)
        ref_raster.append(src.trial_raster(k, src.t - 20, src.t))
    dst = sim(dst_road, np.zeros(B, np.uint64))
    td = table(dst, np.array([B - 1]))
    for _ in range(3):  # the destination's clock is elsewhere
        dst.window(20, td)
    dst.load_trial_state(B - 1, load_trial_state(tmp_path / "trial.npz"))
    for j in range(8):
        dst.window(20, td)
        got = dst.trial_raster(B - 1, dst.t - 20, dst.t)
        assert np.array_equal(got, ref_raster[j]), "window %d" % j
    assert sum(r.sum() for r in ref_raster) > 0
    for name in ("v", "g", "st", "counts"):
        x = getattr(src, name).view(conn.n, src.bstride)[:, k]
        y = getattr(dst, name).view(conn.n, dst.bstride)[:, B - 1]
        assert torch.equal(x, y), name
