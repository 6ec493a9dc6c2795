"""Pinning the oracle (CPU): golden vectors of the reference, published
known answers, closed forms and invariants of the model.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, the new instead of the old conductance, a wrong ring index, a reset to
vReset instead of v - (v - vReset)) fails at least one of them.
"""

from __future__ import annotations

import hashlib

import numpy as np
import pytest

import oracle as O
from flybrain_amd import connectome as C


def fp(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


def canon(a):
    return (a + np.float32(0.0)).astype(np.float32)


def fps_of(eng):
    return [fp(eng.v), fp(eng.conductance), fp(eng.spikes), fp(eng.refrac),
            fp(canon(eng.delay_buffer))]


# --------------------------------------------------------------- golden


def test_tiny_masks_matches_reference(golden):
    """600 steps, B=2, masks switched on and off: all five state tensors
    byte-equal to the reference's after every step."""
    z = golden("tiny_masks.npz")
    conn = C.Connectome(int(z["n"]), z["conn_offsets"], z["conn_cols"],
                        z["conn_vals"])
    assert conn.fingerprint() == str(z["conn_fp"])
    n, B = conn.n, int(z["batch"])
    bits = np.unpackbits(z["tape_bits"], axis=-1)[..., :n]
    eng = O.OracleEngine(conn, B, no_refractory=z["no_refractory"])
    sched = {}
    for k, s in enumerate(z["sched_steps"]):
        kw = {}
        for name in ("fire", "out", "inc", "hold", "v_clamp"):
            key = "sched%d_%s" % (k, name)
            if key in z.files:
                kw[name] = float(z[key]) if name == "v_clamp" else z[key]
        sched[int(s)] = kw
    raster = np.unpackbits(z["raster"], axis=-1)[..., :n]
    for t in range(int(z["steps"])):
        if t in sched:
            eng.set_masks(**sched[t])
        eng.step_bits(bits[t])
        assert fps_of(eng) == list(z["fps"][t]), "step %d" % t
        np.testing.assert_array_equal(canon(eng.rec), z["rec"][t])
        np.testing.assert_array_equal(eng.spikes.astype(np.uint8), raster[t])
    assert int(z["negzero_in_rec"]) == 0


def test_small10k_matches_reference(golden):
    z = golden("small10k.npz")
    conn = C.small()
    assert conn.fingerprint() == str(z["conn_fp"])
    eng = O.OracleEngine(conn, 1, no_refractory=z["driven"])
    sp_t, sp_i = [], []
    for t in range(int(z["steps"])):
        b = np.zeros((1, conn.n), np.uint8)
        b[0, z["tape_i"][z["tape_t"] == t]] = 1
        eng.step_bits(b)
        assert fps_of(eng) == list(z["fps"][t]), "step %d" % t
        idx = np.nonzero(eng.spikes[0])[0]
        sp_t.append(np.full(idx.size, t))
        sp_i.append(idx)
    np.testing.assert_array_equal(np.concatenate(sp_t), z["spike_t"])
    np.testing.assert_array_equal(np.concatenate(sp_i), z["spike_i"])


# --------------------------------------------------------------- Philox


def test_philox_known_answers():
    """Random123 kat_vectors for philox4x32-10."""
    kat = [
        ([0, 0, 0, 0], [0, 0],
         [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2,
         [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
         [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, want in kat:
        assert list(O.philox(ctr, key)) == want


def test_bernoulli_rate():
    """Drive stream: kick frequency matches p within 5 sigma."""
    p = np.float32(150.0) * np.float32(1e-4)
    k = sum(O.draw(7, 0, i % 50, i // 50, p) for i in range(200_000))
    mu = 200_000 * float(p)
    assert abs(k - mu) < 5 * np.sqrt(mu)


def test_uniform_grid_is_the_draw():
    """The grid helper returns exactly the 24-bit uniforms oracle_draw
    compares: draw(p) is u < p * 2^24, so p = u * 2^-24 gives no kick and
    p = (u + 1) * 2^-24 a kick (pins the comparison and its scale)."""
    u = O.uniform_grid(99, 3, 10, 6, 37, 9)
    for k in range(9):
        for i in range(6):
            x = int(u[k, i])
            assert O.draw(99, 3, 10 + i, 37 + k, np.float32(x / 2**24)) == 0
            assert O.draw(99, 3, 10 + i, 37 + k,
                          np.float32((x + 1) / 2**24)) == 1


def _corr(a, b):
    a = a.astype(np.float64).ravel() - a.mean()
    b = b.astype(np.float64).ravel() - b.mean()
    return float((a * b).sum() / np.sqrt((a * a).sum() * (b * b).sum()))


def stream_checks(u, p=0.015):
    """Uniformity and independence of a [steps, neurons] grid of 24-bit
    uniforms: the mean, lag-1 correlations across neighbouring steps
    (within and across the 4-word Philox blocks) and neighbouring neurons,
    and the kick rate and the rate of kicks at two consecutive steps of a
    Bernoulli(p) drive, each within 5 standard errors."""
    x = u.astype(np.float64) / 2**24
    N = x.size
    assert abs(x.mean() - 0.5) < 5 * np.sqrt(1 / 12 / N)
    lim = 5 / np.sqrt(N)
    for lag_t in (1, 2, 3, 4):
        assert abs(_corr(x[lag_t:], x[:-lag_t])) < lim, "step lag %d" % lag_t
    assert abs(_corr(x[:, 1:], x[:, :-1])) < lim, "neuron lag 1"
    kick = u < np.uint32(p * 2**24)
    mu = p * N
    assert abs(kick.sum() - mu) < 5 * np.sqrt(mu * (1 - p))
    both = (kick[1:] & kick[:-1]).sum()
    m2 = p * p * kick[1:].size
    assert abs(both - m2) < 5 * np.sqrt(m2)


def test_drive_stream_is_uniform_and_independent():
    """The Philox drive stream over 1,500 neurons x 2,000 steps of one
    trial, and across trials and seeds (a trial id or seed that did not
    enter the counter or key would correlate the grids)."""
    a = O.uniform_grid(7, 0, 0, 1500, 0, 2000)
    stream_checks(a)
    b = O.uniform_grid(7, 1, 0, 1500, 0, 2000)      # another trial
    c = O.uniform_grid(8, 0, 0, 1500, 0, 2000)      # another seed
    lim = 5 / np.sqrt(a.size)
    assert abs(_corr(a, b)) < lim and abs(_corr(a, c)) < lim
    assert abs(_corr(b, c)) < lim


def test_stream_checks_catch_a_reused_word():
    """The checks above have teeth: a stream that reused one Philox word
    for all four steps of a block (a dropped ``t & 3``), or one that
    ignored the neuron, fails them."""
    a = O.uniform_grid(7, 0, 0, 1500, 0, 2000)
    reused = a[(np.arange(a.shape[0]) // 4) * 4]
    with pytest.raises(AssertionError):
        stream_checks(reused)
    same_neuron = np.repeat(a[:, :1], a.shape[1], axis=1)
    with pytest.raises(AssertionError):
        stream_checks(same_neuron)


def test_run_counts_equals_the_stepped_oracle():
    """oracle_run_counts (the C loop the long statistical comparisons use)
    gives the spike counts of OracleEngine stepped one step at a time on
    the same tape."""
    conn = C.synthetic(900, None, mean_out=50, count_exp=1.3, n_groups=3,
                       seed=2)
    rng = np.random.default_rng(3)
    drv = rng.choice(conn.n, 25, replace=False)
    steps = 400
    kicks = rng.random((steps, drv.size)) < 0.08
    t_, j = np.nonzero(kicks)
    ptr = np.zeros(steps + 1, np.int64)
    np.cumsum(np.bincount(t_, minlength=steps), out=ptr[1:])
    got = O.run_counts(conn, steps, ptr, drv[j], no_refractory=drv)
    ora = O.OracleEngine(conn, 1, no_refractory=drv)
    bits = np.zeros((1, conn.n), np.uint8)
    want = np.zeros(conn.n, np.int64)
    for t in range(steps):
        bits[:] = 0
        bits[0, drv[kicks[t]]] = 1
        want += ora.step_bits(bits)[0].astype(np.int64)
    assert want.sum() > kicks.sum()  # recurrent spikes beyond the kicks
    np.testing.assert_array_equal(got, want)


# --------------------------------------------------------------- closed forms


def two_neurons(count):
    """Neuron 0 -> neuron 1 with ``count`` synapses; nothing else."""
    return C.from_edges(3, [0], [1], [count])


def kick_at(eng, steps, who, when):
    out = []
    for t in range(steps):
        bits = np.zeros((1, eng.n), np.uint8)
        if t in when:
            bits[0, who] = 1
        eng.step_bits(bits)
        out.append((eng.v.copy(), eng.conductance.copy(), eng.spikes.copy(),
                    eng.refrac.copy()))
    return out


def test_isolated_neuron_stays_at_rest():
    eng = O.OracleEngine(two_neurons(5), 1)
    for _ in range(200):
        eng.step_bits(np.zeros((1, 3), np.uint8))
    assert np.all(eng.v == np.float32(-52.0))
    assert np.all(eng.conductance == 0.0)
    assert not eng.spikes.any()


def test_one_kick_spikes_and_resets_exactly():
    """-52 + 68.75 = 16.75, leak to 16.40625, spike, v - (v + 52) = -52."""
    eng = O.OracleEngine(two_neurons(5), 1, no_refractory=[0])
    bits = np.zeros((1, 3), np.uint8)
    bits[0, 0] = 1
    eng.step_bits(bits)
    assert eng.spikes[0, 0] == 1.0
    assert eng.v[0, 0] == np.float32(-52.0)
    assert eng.spikes[0, 1] == 0.0


def test_delay_is_20_steps_to_g_and_21_to_v():
    c = 40
    eng = O.OracleEngine(two_neurons(c), 1, no_refractory=[0])
    hist = kick_at(eng, 30, 0, {3})
    g1 = [h[1][0, 1] for h in hist]
    v1 = [h[0][0, 1] for h in hist]
    assert all(g == 0.0 for g in g1[:23])
    assert g1[23] == np.float32(0.275) * np.float32(c)
    assert all(v == np.float32(-52.0) for v in v1[:24])
    assert v1[24] > np.float32(-52.0)
    # conductance decays by exactly one float32 multiply per step
    assert g1[24] == g1[23] * np.float32(0.98)
    assert g1[25] == g1[24] * np.float32(0.98)


def test_refractory_gate_closes_22_steps():
    """After a spike at t, input arriving at t+1 .. t+22 is dropped and
    input at t+23 is taken (refrac counts 0, 1, ... from t+1; gate r >= 22)."""
    for arrive_after, taken in ((22, False), (23, True)):
        conn = C.from_edges(3, [2], [1], [40])
        eng = O.OracleEngine(conn, 1, no_refractory=[2])
        t_spike = 25
        kick2 = t_spike + arrive_after - 20  # 2 -> 1 lands 20 steps later
        gs, sp = [], []
        for t in range(t_spike + 30):
            bits = np.zeros((1, 3), np.uint8)
            if t == t_spike:
                bits[0, 1] = 1  # a direct kick makes neuron 1 fire
            if t == kick2:
                bits[0, 2] = 1
            eng.step_bits(bits)
            gs.append(eng.conductance[0, 1])
            sp.append(eng.spikes[0, 1])
        assert np.nonzero(sp)[0].tolist() == [t_spike]
        arrive = kick2 + 20
        assert (gs[arrive] != gs[arrive - 1] * np.float32(0.98)) == taken


def tail_length(mask):
    """Steps during which neuron 1 still receives from neuron 0 after the
    mask is put on neuron 0 (neuron 0 kicked, and firing, every step;
    neuron 1 held so it never fires and its refractory gate stays open)."""
    conn = two_neurons(3)
    eng = O.OracleEngine(conn, 1, no_refractory=[0])
    eng.set_masks(hold=[1])
    on = 60
    received = []
    for t in range(120):
        if t == on:
            eng.set_masks(**{mask: [0]})
        g_prev = eng.conductance[0, 1]
        bits = np.zeros((1, 3), np.uint8)
        bits[0, 0] = 1
        eng.step_bits(bits)
        received.append(eng.conductance[0, 1] != g_prev * np.float32(0.98))
    last = max(t for t, r in enumerate(received) if r)
    return last - on + 1


def test_mask_tails_out_19_fire_20():
    """lab/flybrain/sim.py:886-897: m_out tail 19 steps, m_fire 20."""
    assert tail_length("out") == 19
    assert tail_length("fire") == 20


def test_inc_mask_is_instant():
    conn = two_neurons(3)
    eng = O.OracleEngine(conn, 1, no_refractory=[0])
    for t in range(60):
        if t == 40:
            eng.set_masks(inc=[1])
        g_prev = eng.conductance[0, 1]
        bits = np.zeros((1, 3), np.uint8)
        bits[0, 0] = 1
        eng.step_bits(bits)
        if t >= 40:
            assert eng.conductance[0, 1] == g_prev * np.float32(0.98)


def test_hold_clamps_exactly():
    conn = two_neurons(3)
    eng = O.OracleEngine(conn, 1, no_refractory=[0])
    eng.set_masks(hold=[1], v_clamp=-61.25)
    for t in range(80):
        bits = np.zeros((1, 3), np.uint8)
        bits[0, 0] = 1
        eng.step_bits(bits)
        assert eng.v[0, 1] == np.float32(-61.25)
        assert eng.spikes[0, 1] == 0.0


def test_all_ones_masks_are_byte_neutral(golden):
    z = golden("tiny_masks.npz")
    conn = C.Connectome(int(z["n"]), z["conn_offsets"], z["conn_cols"],
                        z["conn_vals"])
    n, B = conn.n, int(z["batch"])
    bits = np.unpackbits(z["tape_bits"], axis=-1)[..., :n]
    a = O.OracleEngine(conn, B, no_refractory=z["no_refractory"])
    b = O.OracleEngine(conn, B, no_refractory=z["no_refractory"])
    b.set_masks(fire=[], out=[], inc=[], hold=[])
    for t in range(100):
        a.step_bits(bits[t])
        b.step_bits(bits[t])
    assert fps_of(a) == fps_of(b)


def test_scatter_equals_dense_product():
    """The oracle's event scatter equals a dense int64 W @ s."""
    conn = C.synthetic(300, None, mean_out=30, n_groups=3, seed=4)
    w = conn.dense_counts()
    rng = np.random.default_rng(1)
    eng = O.OracleEngine(conn, 2)
    eng.spikes[:] = (rng.random((2, conn.n)) < 0.3).astype(np.float32)
    eng.step_bits(np.zeros((2, conn.n), np.uint8))
    s = (rng.random((2, conn.n)) < 0.3)  # consumed above; recompute below
    del s
    eng2 = O.OracleEngine(conn, 2)
    sp = (np.random.default_rng(2).random((2, conn.n)) < 0.3).astype(np.float32)
    eng2.spikes[:] = sp
    eng2.step_bits(np.zeros((2, conn.n), np.uint8))
    np.testing.assert_array_equal(eng2.icount, (w @ sp.T.astype(np.int64)).T)


def test_rate_limit_is_loud():
    from flybrain_amd.device import RateTooHigh, check_rates_hz

    check_rates_hz(9999.0, "x")
    check_rates_hz(10000.0, "x")
    with pytest.raises(RateTooHigh):
        check_rates_hz(10001.0, "x")
    with pytest.raises(RateTooHigh):
        check_rates_hz([-1.0, 5.0], "x")


def test_torch_restatement_matches_reference(golden):
    """The torch-CPU restatement (the CPU baseline) reproduces the golden
    fingerprints on the 10k workload for the first 300 steps."""
    import ref_torch

    z = golden("small10k.npz")
    conn = C.small()
    eng = ref_torch.TorchRef(conn, 1, no_refractory=z["driven"])
    for t in range(300):
        b = np.zeros((1, conn.n), np.uint8)
        b[0, z["tape_i"][z["tape_t"] == t]] = 1
        eng.step(b)
        got = [fp(eng.v.numpy()), fp(eng.conductance.numpy()),
               fp(eng.spikes.numpy()), fp(eng.refrac.numpy()),
               fp(canon(eng.delay_buffer.numpy()))]
        assert got == list(z["fps"][t]), "step %d" % t
