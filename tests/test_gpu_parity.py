"""Parity of the HIP path with the oracle and the reference's golden vectors.

Bar (integer/byte work bit-exact; float32 state within 1e-4 mV, measured
here as bit-exact because every float op keeps the reference's rounding):
  * per-step synaptic input (integer synapse counts) identical;
  * spike raster identical;
  * v and g bit-identical, refractory counter = min(reference, 127).
"""

from __future__ import annotations

import hashlib

import numpy as np
import pytest
import torch

import oracle as O
from flybrain_amd import connectome as C
from flybrain_amd.device import DeviceNet, DeviceSim, DriveTable, Tape
from flybrain_amd._lib import OUT_OFF as O_OUT_OFF
from gpu_helpers import (apply_masks_device, compare_state, mask_schedule,
                         state_of, tiny_conn)

pytestmark = pytest.mark.gpu


def fp(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


def tape_of_bits(bits):
    """(steps, neurons, trials) of a [steps, B, n] 0/1 array."""
    t, b, i = np.nonzero(bits)
    return t, i, b


@pytest.fixture(scope="module")
def tiny(golden):
    z = golden("tiny_masks.npz")
    conn = tiny_conn(z)
    n = conn.n
    bits = np.unpackbits(z["tape_bits"], axis=-1)[..., :n]
    return z, conn, bits


@pytest.mark.parametrize("road", ["push", "pull"])
def test_tiny_masks_per_step(tiny, road):
    """W=1 steps, masks switched as in the fixture: every step equals the
    oracle and the reference's own fingerprints of v, g and spikes."""
    z, conn, bits = tiny
    B, steps = int(z["batch"]), int(z["steps"])
    net = DeviceNet(conn)
    sim = DeviceSim(net, B, no_refractory=z["no_refractory"], raster_len=32,
                    road=road)
    t_, i_, b_ = tape_of_bits(bits)
    tape = Tape(sim, t_, i_, b_)
    ora = O.OracleEngine(conn, B, no_refractory=z["no_refractory"])
    sched = mask_schedule(z)
    for t in range(steps):
        if t in sched:
            apply_masks_device(sim, sched[t])
            ora.set_masks(**sched[t])
        sim.window(1, tape)
        ora.step_bits(bits[t])
        v, g, sp, r = state_of(sim)
        assert fp(v) == z["fps"][t][0], "step %d: v fingerprint" % t
        assert fp(g) == z["fps"][t][1], "step %d: g fingerprint" % t
        assert fp(sp) == z["fps"][t][2], "step %d: spikes fingerprint" % t
        compare_state(sim, ora, "step %d" % t)
        got = sim.rec_slot(t).cpu().numpy()
        np.testing.assert_array_equal(got, ora.icount,
                                      err_msg="step %d: synaptic input" % t)


@pytest.mark.parametrize("road", ["push", "pull"])
def test_tiny_masks_windows(tiny, road):
    """Windows of up to 20 steps between mask changes: raster and state at
    every window end equal the oracle; synaptic input equals the golden."""
    z, conn, bits = tiny
    B, steps = int(z["batch"]), int(z["steps"])
    net = DeviceNet(conn)
    sim = DeviceSim(net, B, no_refractory=z["no_refractory"], raster_len=64,
                    road=road)
    t_, i_, b_ = tape_of_bits(bits)
    tape = Tape(sim, t_, i_, b_)
    ora = O.OracleEngine(conn, B, no_refractory=z["no_refractory"])
    sched = mask_schedule(z)
    cuts = sorted(set(range(0, steps, 20)) | set(sched) | {steps})
    rec_golden = z["rec"]
    for a, b in zip(cuts[:-1], cuts[1:]):
        while a < b:
            W = min(20, b - a)
            if a in sched:
                apply_masks_device(sim, sched[a])
                ora.set_masks(**sched[a])
            sim.window(W, tape)
            raster = sim.read_raster(a, a + W)
            for k in range(W):
                ora.step_bits(bits[a + k])
                assert np.array_equal(raster[k], ora.spikes.astype(bool)), \
                    "raster differs at step %d" % (a + k)
            e = a + W - 1
            for u in range(max(a, e - 18), e + 1):
                # the ring still holds rec(u) for the last 19 steps
                want = np.rint(rec_golden[u] / np.float32(0.275))
                np.testing.assert_array_equal(
                    sim.rec_slot(u).cpu().numpy(), want.astype(np.int32),
                    err_msg="synaptic input of step %d" % u)
            compare_state(sim, ora, "window end %d" % e)
            a += W


def test_small10k_config1(golden):
    """10,000 neurons / 1.09M edges, 20 neurons at 150 Hz from the recorded
    tape, 1,000 steps in 20-step windows: raster and final state equal the
    reference's golden vectors."""
    z = golden("small10k.npz")
    conn = C.small()
    assert conn.fingerprint() == str(z["conn_fp"])
    steps = int(z["steps"])
    net = DeviceNet(conn)
    sim = DeviceSim(net, 1, no_refractory=z["driven"], raster_len=1000)
    tape = Tape(sim, z["tape_t"], z["tape_i"])
    for a in range(0, steps, 20):
        sim.window(20, tape)
    raster = sim.read_raster(0, steps)
    t, _, i = np.nonzero(raster)
    assert t.size == z["spike_t"].size
    assert np.array_equal(t, z["spike_t"]) and np.array_equal(i, z["spike_i"])
    v, g, sp, r = state_of(sim)
    assert fp(v) == z["fps"][-1][0]
    assert fp(g) == z["fps"][-1][1]
    np.testing.assert_array_equal(r[0], np.minimum(z["refrac_final"][0], 127))


@pytest.mark.parametrize("batch,road,fused", [(3, "push", False),
                                              (64, "push", False),
                                              (130, "push", False),
                                              (64, "pull", False),
                                              (100, "pull", False),
                                              (130, "pull", False),
                                              (256, "pull", False),
                                              (300, "pull", False),
                                              (600, "pull", False),
                                              (3, "push", True),
                                              (64, "push", True),
                                              (130, "push", True)])
def test_philox_drive_batched(batch, road, fused):
    """Dense-rate Philox drive (Engine.step road) and per-trial drive tables
    against the oracle running the same Philox stream, trial by trial."""
    philox_batched(batch, road, fused)


@pytest.mark.parametrize("batch,tile", [(64, 128), (64, 512), (100, 128),
                                        (256, 128), (256, 512), (600, 128),
                                        (600, 512)])
def test_pull_tiles_match_oracle(batch, tile):
    """Every compiled pull-road workgroup shape (128 / 512 neuron-trials,
    one to eight neurons per tile, or a part of one neuron's trials) gives
    the oracle's raster and state."""
    philox_batched(batch, "pull", False, tile)


@pytest.mark.parametrize("batch,road,refrac", [(64, "pull", 0),
                                                (64, "pull", 3),
                                                (256, "pull", 7),
                                                (64, "pull", 126),
                                                (3, "push", 3)])
def test_refractory_periods_match_oracle(batch, road, refrac):
    """Refractory periods shorter than a window (the gate reopens within
    it), zero, and the longest the engine takes: raster and state equal the
    oracle, on the pull road's plain and flagged waves alike."""
    philox_batched(batch, road, False,
                   params=dict(O.MODEL_PARAMS, tRefrac=refrac * 0.1))


@pytest.mark.parametrize("batch,tile,n_drive", [(64, 256, (1, 3)),
                                                (256, 256, (1, 3)),
                                                (100, 512, (1, 4)),
                                                (600, 128, (1, 2)),
                                                (256, 128, (5, 40))])
def test_pull_gated_matches_oracle(batch, tile, n_drive):
    """Input gating (tiles that no input can reach skip their in-edge
    walk): with a sparse drive most tiles skip most windows, and raster
    and state still equal the oracle; so with a dense one."""
    sim = philox_batched(batch, "pull", False, tile, gate=True,
                         n_drive=n_drive, conn_kw=dict(mean_out=6))
    due = sim.due.cpu().numpy()
    assert (due >= sim.t).any()
    if n_drive[1] <= 4:  # sparse: some tiles had no input to gather
        assert (due < sim.t).any()


def philox_batched(batch, road, fused, tile=None, params=None, gate=False,
                   n_drive=(5, 40), conn_kw=None, alike=False, lists=True):
    """``alike``: every trial the same seed and drive (each spike is then
    one of every trial: the record lists overflow); ``lists=False``: the
    pull road reads the record words only."""
    kw = dict(mean_out=40, count_exp=1.2, n_groups=4, seed=21)
    kw.update(conn_kw or {})
    conn = C.synthetic(512, None, **kw)
    n = conn.n
    rng = np.random.default_rng(batch)
    seeds = rng.integers(0, 2**63, size=batch, dtype=np.uint64)
    drv = [rng.choice(n, rng.integers(*n_drive), replace=False)
           for _ in range(batch)]
    rate = [float(rng.uniform(50, 900)) for _ in range(batch)]
    if alike:
        seeds[:] = seeds[0]
        drv = [drv[0]] * batch
        rate = [rate[0]] * batch
    net = DeviceNet(conn)
    sim = DeviceSim(net, batch, params=params, seeds=seeds, raster_len=40,
                    road=road, pull_tile=tile, gate=gate)
    if not lists:
        sim.rec_ids = None
        sim._sync_struct()
    if params:
        assert sim.c["refrac_steps"] == round(params["tRefrac"] / 0.1)
    ne = np.concatenate(drv)
    tr = np.concatenate([np.full(len(d), b) for b, d in enumerate(drv)])
    sim.exempt_pairs(ne, tr)  # refractory exemption per trial
    table = DriveTable(sim, ne, tr, tr, np.array(rate, np.float32))
    if fused:  # the update launches draw the table's kicks themselves
        sim.fuse_drive(table)
    ora = O.OracleEngine(conn, batch, params=params, seeds=seeds,
                         no_refractory_trials=drv)
    rates = np.zeros((batch, n), np.float32)
    for b, d in enumerate(drv):
        rates[b, d] = rate[b]
    steps = 400
    for a in range(0, steps, 20):
        sim.window(20, None if fused else table)
        raster = sim.read_raster(a, a + 20)
        for k in range(20):
            ora.step_rates(rates)
            assert np.array_equal(raster[k], ora.spikes.astype(bool)), \
                "raster differs at step %d" % (a + k)
    compare_state(sim, ora, "end")
    return sim


@pytest.mark.parametrize("batch,tile", [(256, 256), (64, 128), (64, 256)])
def test_windows_from_rest_into_a_burst(batch, tile):
    """The pull road's recurrence paths against the oracle, each made to
    run: 60 steps at rest (every input gate open: waves take the window
    without gate and reset, or none at all where nothing moves), then a
    tape burst from step 65 into a third of the trials -- the driven waves
    (no refractoriness, kicks) take the kick recurrence, and waves of their
    targets that start the window open and fire in it run it again with
    gate and reset; the next windows hold refractory lanes."""
    conn = C.synthetic(512, None, mean_out=40, count_exp=1.2, n_groups=4,
                       seed=5)
    n, steps = conn.n, 140
    rng = np.random.default_rng(batch + tile)
    drv = rng.choice(n, 24, replace=False)
    trials = np.arange(batch)[rng.random(batch) < 1 / 3]
    bits = np.zeros((steps, batch, n), np.uint8)
    burst = rng.random((steps - 65, trials.size, drv.size)) < 0.6
    bits[65:, trials[:, None], drv[None, :]] = burst
    t_, i_, b_ = tape_of_bits(bits)
    sim = DeviceSim(DeviceNet(conn), batch, no_refractory=drv,
                    raster_len=steps, road="pull", pull_tile=tile)
    ora = O.OracleEngine(conn, batch, no_refractory=drv)
    tape = Tape(sim, t_, i_, b_)
    for a in range(0, steps, 20):
        sim.window(20, tape)
        if a == 40:
            compare_state(sim, ora_run(ora, bits, 0, 60), "rest")
    raster = sim.read_raster(0, steps)
    ora_run(ora, bits, 60, steps)
    ora2 = O.OracleEngine(conn, batch, no_refractory=drv)
    fired = np.zeros(n, bool)
    for k in range(steps):
        ora2.step_bits(bits[k])
        assert np.array_equal(raster[k], ora2.spikes.astype(bool)), \
            "raster differs at step %d" % k
        fired |= ora2.spikes.any(0)
    compare_state(sim, ora, "end")
    assert not raster[20:60].any(), "the network was not at rest"
    undriven = np.setdiff1d(np.arange(n), drv)
    assert fired[undriven].sum() >= 5, "the burst reached no target"


def ora_run(ora, bits, t0, t1):
    for k in range(t0, t1):
        ora.step_bits(bits[k])
    return ora


@pytest.mark.parametrize("batch", [64, 100, 200, 600])
def test_pull_road_uneven_windows_and_out_masks(batch):
    """Pull road with a partial last tile (509 neurons), windows of 1..20
    steps (records of one delay horizon span several windows) and m_out
    switched between windows on spiking neurons: raster, state and the
    in-flight input equal the oracle, with and without input gating (the
    m_out re-send stamps the targets it re-enables); the push road gives
    the same."""
    conn = C.synthetic(509, None, mean_out=45, count_exp=1.15, n_groups=4,
                       seed=23)
    n = conn.n
    rng = np.random.default_rng(batch + 7)
    seeds = rng.integers(0, 2**63, size=batch, dtype=np.uint64)
    drv = [rng.choice(n, rng.integers(10, 60), replace=False)
           for _ in range(batch)]
    rate = rng.uniform(100, 1500, size=batch).astype(np.float32)
    ne = np.concatenate(drv)
    tr = np.concatenate([np.full(len(d), b) for b, d in enumerate(drv)])
    net = DeviceNet(conn)
    sims = {}
    for road in ("pull", "pull_gated", "push"):
        sim = DeviceSim(net, batch, seeds=seeds, raster_len=40,
                        road=road.split("_")[0], gate=road == "pull_gated")
        sim.exempt_pairs(ne, tr)
        sims[road] = (sim, DriveTable(sim, ne, tr, tr, rate))
    ora = O.OracleEngine(conn, batch, seeds=seeds, no_refractory_trials=drv)
    rates = np.zeros((batch, n), np.float32)
    for b, d in enumerate(drv):
        rates[b, d] = rate[b]
    hubs = np.argsort(-conn.degree())[:40]
    plan = [20, 7, 13, 1, 19, 20, 5, 20, 11, 9, 20, 3, 17, 20, 20, 2, 18]
    masks = {2: hubs[:25], 5: hubs[10:], 9: [], 12: hubs[::3]}
    t = 0
    for k, W in enumerate(plan):
        if k in masks:
            for sim, _ in sims.values():
                sim.set_flag(O_OUT_OFF, masks[k])
            ora.set_masks(out=list(masks[k]))
        for sim, table in sims.values():
            sim.window(W, table)
        got = {r: s.read_raster(t, t + W) for r, (s, _) in sims.items()}
        for j in range(W):
            ora.step_rates(rates)
            for r in sims:
                want = ora.spikes.astype(bool)
                if not np.array_equal(got[r][j], want):
                    bt, nn = np.nonzero(got[r][j] != want)
                    driven = [bool(rates[b, i] > 0) for b, i in zip(bt, nn)]
                    pytest.fail(
                        "%s road: raster differs at step %d: %d neuron-trials"
                        ", first (trial, neuron, got, driven) %s"
                        % (r, t + j, bt.size,
                           list(zip(bt[:8].tolist(), nn[:8].tolist(),
                                    got[r][j][bt[:8], nn[:8]].tolist(),
                                    driven[:8]))))
        t += W
        for r, (s, _) in sims.items():
            compare_state(s, ora, "%s road, step %d" % (r, t))
        img = {r: s.ring_image().cpu().numpy() for r, (s, _) in sims.items()}
        for r in ("pull", "pull_gated"):
            np.testing.assert_array_equal(
                img[r].reshape(20, n, -1)[:, :, :batch],
                img["push"].reshape(20, n, -1)[:, :, :batch],
                err_msg="%s: input in flight at step %d" % (r, t))
    assert ora.spikes.any()


def test_rates_drive_matches_table():
    """fis_drive_rates (dense [B, n] rates, Engine.step road) produces the
    same kicks as a drive table of the same rates."""
    conn = C.synthetic(300, None, mean_out=30, n_groups=0, seed=3)
    n, B = conn.n, 2
    net = DeviceNet(conn)
    a = DeviceSim(net, B, seeds=[11, 12], raster_len=20)
    b = DeviceSim(net, B, seeds=[11, 12], raster_len=20)
    rates = torch.zeros(B, n, device="cuda")
    idx = np.arange(0, n, 7)
    rates[0, idx] = 2000.0
    rates[1, idx[::2]] = 7000.0
    ne = np.concatenate([idx, idx[::2]])
    tr = np.concatenate([np.zeros(idx.size), np.ones(idx[::2].size)])
    table = DriveTable(a, ne, tr.astype(np.int64), np.arange(ne.size),
                       rates[tr.astype(np.int64), ne].cpu().numpy())
    for t in range(60):
        a.window(1, rates)
        b.window(1, table)
        assert np.array_equal(a.read_raster(t, t + 1), b.read_raster(t, t + 1))
    torch.testing.assert_close(a.v, b.v, rtol=0, atol=0)


def test_ev_spmv_binds_by_the_reference_names():
    """libfis.so stands where the reference's ev_spmv DLL stood: it is
    opened with a fresh ctypes.CDLL, checked with ev_spmv_abi() == 3 and its
    ev_spmv_i32 / ev_spmv_i64 are bound by those names with the argtypes of
    lab/flybrain/evspmv.py:886-903, then called the way Events.mm calls them
    (evspmv.py:1115-1138: out zeroed, raw data_ptr() ints, the current
    stream's handle).  The result equals a dense float64 product on the
    CPU."""
    import ctypes

    from flybrain_amd import _lib as L

    dll = ctypes.CDLL(str(L.LIB_PATH))
    dll.ev_spmv_abi.restype = ctypes.c_int
    dll.ev_spmv_abi.argtypes = []
    assert int(dll.ev_spmv_abi()) == 3
    call = {}
    for key_t, name in (("int32", "ev_spmv_i32"), ("int64", "ev_spmv_i64")):
        fn = getattr(dll, name)
        fn.restype = ctypes.c_int
        fn.argtypes = ([ctypes.c_void_p] * 6
                       + [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                          ctypes.c_void_p])
        call[key_t] = fn

    conn = C.synthetic(700, None, mean_out=50, n_groups=0, seed=5)
    n, B = conn.n, 3
    rng = np.random.default_rng(0)
    sp = (rng.random((B, n)) < 0.2).astype(np.float32)
    mo = (rng.random((B, n)) < 0.9).astype(np.float32)
    w = conn.dense_counts().astype(np.float64)  # [post, pre]
    want = (w @ (sp * mo).T.astype(np.float64)).T
    for key_t, dt in (("int32", torch.int32), ("int64", torch.int64)):
        crow = torch.from_numpy(conn.offsets).to("cuda", dt)
        col = torch.from_numpy(conn.cols.astype(np.int64)).to("cuda", dt)
        val = torch.from_numpy(conn.vals.astype(np.float32)).cuda()
        spt, mot = torch.from_numpy(sp).cuda(), torch.from_numpy(mo).cuda()
        out = torch.full((B, n), 7.0, device="cuda")
        for _ in range(2):  # the persistent out, zeroed before each launch
            out.zero_()
            rc = call[key_t](crow.data_ptr(), col.data_ptr(), val.data_ptr(),
                             spt.data_ptr(), mot.data_ptr(), out.data_ptr(),
                             n, B, 128,
                             torch.cuda.current_stream().cuda_stream)
            assert rc == 0
            np.testing.assert_array_equal(out.cpu().numpy(),
                                          want.astype(np.float32))


@pytest.mark.parametrize("batch", [1, 3, 32])
def test_one_launch_push_window_equals_two(batch):
    """Below 64 trials the push road runs update and scatter as one launch
    (fis_update_scatter: spikes added straight into ring slots the window
    does not read).  The two-launch form (fis_update + fis_scatter, event
    records) gives the same raster, state and ring window by window, over
    windows of 1..20 steps, and both equal the oracle."""
    conn = C.synthetic(512, None, mean_out=40, count_exp=1.2, n_groups=4,
                       seed=29)
    n = conn.n
    rng = np.random.default_rng(batch + 11)
    seeds = rng.integers(0, 2**63, size=batch, dtype=np.uint64)
    drv = [rng.choice(n, rng.integers(5, 40), replace=False)
           for _ in range(batch)]
    rate = rng.uniform(100, 900, size=batch).astype(np.float32)
    net = DeviceNet(conn)
    ne = np.concatenate(drv)
    tr = np.concatenate([np.full(len(d), b) for b, d in enumerate(drv)])
    sims = []
    for _ in range(2):
        sim = DeviceSim(net, batch, seeds=seeds, raster_len=40, road="push")
        sim.exempt_pairs(ne, tr)
        sims.append((sim, DriveTable(sim, ne, tr, tr, rate)))
    (one, t_one), (two, t_two) = sims
    ora = O.OracleEngine(conn, batch, seeds=seeds, no_refractory_trials=drv)
    rates = np.zeros((batch, n), np.float32)
    for b, d in enumerate(drv):
        rates[b, d] = rate[b]
    a = 0
    for W in (20, 7, 13, 1, 19, 20, 20):
        one.drive_table(t_one, W)
        one.update_scatter(W)
        two.drive_table(t_two, W)
        two.update(W)
        two.scatter()
        r_one = one.read_raster(a, a + W)
        assert np.array_equal(r_one, two.read_raster(a, a + W)), \
            "rasters differ in the window at %d" % a
        for k in range(W):
            ora.step_rates(rates)
            assert np.array_equal(r_one[k], ora.spikes.astype(bool)), \
                "raster differs from the oracle at step %d" % (a + k)
        compare_state(one, ora, "one launch, window end %d" % (a + W - 1))
        compare_state(two, ora, "two launches, window end %d" % (a + W - 1))
        assert torch.equal(one.ring, two.ring), "rings differ at %d" % a
        a += W


@pytest.mark.parametrize("road,batch", [("push", 64), ("push", 3),
                                        ("push", 130), ("pull", 64),
                                        ("pull", 200)])
def test_outputs_stay_inside_their_buffers(road, batch):
    """509 neurons: the last workgroup has waves past the last neuron.
    Guard words placed right after every buffer a launch writes per wave or
    per workgroup (events, event counts, raster, spike records, state) are
    still intact after windows on either road."""
    conn = C.synthetic(509, None, mean_out=45, count_exp=1.15, n_groups=4,
                       seed=23)
    n = conn.n
    rng = np.random.default_rng(batch)
    drv = rng.choice(n, 120, replace=False)
    net = DeviceNet(conn)
    sim = DeviceSim(net, batch, seeds=rng.integers(0, 2**63, size=batch,
                                                   dtype=np.uint64),
                    raster_len=20, road=road, counts=True)
    sentinel = -0x5A5A5A5A5A5A5A5B
    guards = {}
    names = ["events", "ev_count", "raster", "v", "g", "sbits", "counts",
             "ring"]
    if road == "pull":
        names += ["rec", "act", "rec_ids"]
    for name in names:
        t = getattr(sim, name)
        nbytes = t.numel() * t.element_size()
        words = (nbytes + 7) // 8 + 512  # 4 KB of guard after the data
        big = torch.full((words,), sentinel, dtype=torch.int64,
                         device=t.device)
        body = big.view(torch.uint8)[:nbytes].view(t.dtype)
        body.copy_(t)
        setattr(sim, name, body)
        guards[name] = (big, (nbytes + 7) // 8)
    sim._sync_struct()
    ne = np.repeat(drv, batch)
    tr = np.tile(np.arange(batch), drv.size)
    table = DriveTable(sim, ne, tr, tr, np.full(batch, 900.0, np.float32))
    for W in (20, 20, 7, 13, 20):
        sim.window(W, table)
    torch.cuda.synchronize()
    assert sim.read_raster(sim.t - 20, sim.t).any()
    for name, (big, start) in guards.items():
        tail = big[start:].cpu().numpy()  # words after the last data byte
        assert (tail == sentinel).all(), name


@pytest.mark.parametrize("batch", [1, 3, 64, 130])
def test_dense_rates_drawn_in_the_update_equal_the_drive_launch(batch):
    """Push road: kicks from dense rates drawn by the update launch itself
    (fis_sim.drive_dense) equal those of the separate fis_drive_rates
    launch, over windows of 1..20 steps with the rates changed between
    windows (zero, sparse and dense rows)."""
    conn = C.synthetic(300, None, mean_out=30, count_exp=1.2, n_groups=2,
                       seed=13)
    n = conn.n
    rng = np.random.default_rng(batch)
    seeds = rng.integers(0, 2**63, size=batch, dtype=np.uint64)
    net = DeviceNet(conn)
    fused = DeviceSim(net, batch, seeds=seeds, raster_len=40, road="push")
    split = DeviceSim(net, batch, seeds=seeds, raster_len=40, road="push")
    t = 0
    for k, W in enumerate([20, 1, 7, 20, 13, 1, 20, 3]):
        rates = torch.zeros(batch, n, device="cuda")
        if k % 3 != 2:  # every third window undriven
            idx = torch.from_numpy(rng.choice(n, 40, replace=False)).cuda()
            rates[:, idx] = float(rng.uniform(200, 4000))
            rates[0, ::17] = 800.0
        fused.step_rates(rates, W)
        split.drive_rates(rates, W)
        split.step_launch(W)
        assert np.array_equal(fused.read_raster(t, t + W),
                              split.read_raster(t, t + W)), "window %d" % k
        t += W
    torch.testing.assert_close(fused.v, split.v, rtol=0, atol=0)
    torch.testing.assert_close(fused.g, split.g, rtol=0, atol=0)
    assert fused.read_raster(t - 20, t).any()


@pytest.mark.parametrize("batch,tile", [(64, 256), (256, 256), (128, 128)])
def test_calm_tiles_wake_up_exactly(batch, tile):
    """Gated pull road: (neuron, part)s that settle at a fixed point go
    calm and skip windows unread; then a drive switches on at step 300
    (kicks reach calm neurons: the drive launch ends their calm) and
    its spikes reach calm targets (input due ends the skip).  Raster and
    state equal the oracle throughout; calm entries were set and later
    cleared; every record list under a set act bit names the trials of its
    record words (calm tiles leave their own stale, under clear act
    bits)."""
    conn = C.synthetic(640, None, mean_out=8, count_exp=1.2, n_groups=4,
                       seed=31)
    n = conn.n
    rng = np.random.default_rng(batch + tile)
    seeds = rng.integers(0, 2**63, size=batch, dtype=np.uint64)
    # an early drive in four trials only: the parts of neurons it never
    # reaches stay at rest in every trial and go calm
    early = [rng.choice(n, 2 if b < 4 else 0, replace=False)
             for b in range(batch)]
    late = [rng.choice(n, 6, replace=False) for _ in range(batch)]
    drv = [np.unique(np.concatenate([a, b])) for a, b in zip(early, late)]
    net = DeviceNet(conn)
    sim = DeviceSim(net, batch, seeds=seeds, raster_len=40, road="pull",
                    pull_tile=tile, gate=True)
    # parts of 128 and 256 trials also keep record lists
    lists = min(sim.bstride, tile) // 64 >= 2
    ne = np.concatenate(drv)
    tr = np.concatenate([np.full(len(d), b) for b, d in enumerate(drv)])
    sim.exempt_pairs(ne, tr)
    # group 0: the early neurons at 60 Hz throughout; group 1: the late
    # ones, off for 300 steps, then 700 Hz
    grp = np.concatenate([np.where(np.isin(d, a), 0, 1)
                          for d, a in zip(drv, early)])
    table = DriveTable(sim, ne, tr, grp,
                       np.array([[60.0, 0.0], [60.0, 0.0], [60.0, 0.0],
                                 [60.0, 700.0]], np.float32), seg_steps=100)
    ora = O.OracleEngine(conn, batch, seeds=seeds, no_refractory_trials=drv)
    calm_seen = 0
    steps = 800
    for a in range(0, steps, 20):
        sim.window(20, table)
        raster = sim.read_raster(a, a + 20)
        rates = np.zeros((batch, n), np.float32)
        for k in range(20):
            t = a + k
            for b, d in enumerate(drv):
                on = np.isin(d, early[b])
                rates[b, d[on]] = 60.0
                rates[b, d[~on]] = 700.0 if t >= 300 else 0.0
            ora.step_rates(rates)
            assert np.array_equal(raster[k], ora.spikes.astype(bool)), \
                "raster differs at step %d" % t
        if a + 20 == 300:
            calm_seen = int(sim.calm.sum().item())
            if lists:  # calm tiles keep stale words under clear act bits
                check_record_lists(sim)
    compare_state(sim, ora, "end")
    assert calm_seen > 0, "no (neuron, part) went calm before the drive"
    assert int(sim.calm.sum().item()) < calm_seen
    if lists:
        checked, _ = check_record_lists(sim)
        assert checked > 0


def check_record_lists(sim):
    """Every valid record list (act bit set, count <= 15) names exactly the
    trials whose record-word bits are set; returns (lists checked, lists
    that send readers to the words)."""
    n, bs = sim.n, sim.bstride
    tpn = min(bs, sim.pull_tile)
    parts, wpt, words = bs // tpn, tpn // 64, bs // 64
    rec = sim.rec.cpu().numpy().view(np.uint64).reshape(n, 40, words)
    act = sim.act.cpu().numpy().view(np.uint64).reshape(n, parts)
    ids = sim.rec_ids.cpu().numpy().reshape(n, parts, 40, 16)
    checked = to_words = 0
    for slot in range(40):
        valid = ((act >> np.uint64(slot)) & np.uint64(1)).astype(bool)
        for part in range(parts):
            for i in np.nonzero(valid[:, part])[0]:
                w = rec[i, slot, part * wpt:(part + 1) * wpt]
                want = sorted(q * 64 + b for q in range(wpt) for b in range(64)
                              if (int(w[q]) >> b) & 1)
                cnt = int(ids[i, part, slot, 0])
                if cnt == 0xFF:
                    to_words += 1
                    continue
                got = sorted(ids[i, part, slot, 1:1 + cnt].tolist())
                assert got == want, (i, part, slot, got, want)
                checked += 1
    return checked, to_words


@pytest.mark.parametrize("batch,tile", [(128, 128), (256, 256), (256, 128),
                                        (300, 256), (600, 256)])
def test_record_lists_name_the_record_words(batch, tile):
    """Parts of 128 and 256 trials read short trial lists in place of the
    record words: after 400 steps against the oracle, every valid list
    names exactly the trials its words hold."""
    sim = philox_batched(batch, "pull", False, tile)
    checked, _ = check_record_lists(sim)
    assert checked > 0


@pytest.mark.parametrize("batch,tile", [(256, 256), (128, 128), (600, 256)])
def test_record_lists_overflow_to_the_words(batch, tile):
    """Every trial alike: each spike is one of all the part's trials, no
    list holds them and every reader takes the words; raster and state
    still equal the oracle."""
    sim = philox_batched(batch, "pull", False, tile, alike=True)
    checked, to_words = check_record_lists(sim)
    assert to_words > 0


@pytest.mark.parametrize("batch,tile", [(256, 256), (300, 256)])
def test_pull_road_without_record_lists(batch, tile):
    """rec_ids is optional: without it the pull road reads the record
    words, with the same raster and state."""
    philox_batched(batch, "pull", False, tile, lists=False)
