"""Run results on disk (CPU): save() writes the reference's files --
meta.json, rates.parquet (index = neuron, rate_hz, n_spikes), bins.npz --
and load() rebuilds the Result (lab/flybrain/sim.py:1120-1174)."""

from __future__ import annotations

import json

import numpy as np
import pandas as pd
import pytest

from flybrain_amd import sim as S


def _result(n=300, n_bins=25, trials=2, seed=0):
    rng = np.random.default_rng(seed)
    counts = rng.poisson(3.0, size=n).astype(np.int64)
    counts[rng.choice(n, 40, replace=False)] = 0  # silent neurons
    bins = rng.poisson(0.4, size=(n_bins, n)).astype(np.int16)
    bins[:, counts == 0] = 0
    cfg = S.RunConfig(t_run_s=0.25, bin_ms=10.0, n_trials=trials, seed=7)
    stim = [S.Stimulus(indices=np.arange(5, 15), rate_hz=120.0, t_on_ms=0.0,
                       t_off_ms=100.0, name="a")]
    return S.Result(counts, bins, cfg, stim, np.array([3, 4], np.int64),
                    np.arange(n), {"wall_s": 1.5})


def test_save_writes_the_reference_files(tmp_path):
    res = _result()
    S.save(res, tmp_path, min_rate=0.5)
    assert {p.name for p in tmp_path.iterdir()} == {
        "meta.json", "rates.parquet", "bins.npz"}
    df = pd.read_parquet(tmp_path / "rates.parquet")
    keep = res.active(0.5)
    assert sorted(df.index.tolist()) == keep.tolist()
    assert list(df.columns) == ["rate_hz", "n_spikes"]
    assert (np.diff(df["rate_hz"].to_numpy()) <= 0).all()  # by rate, desc
    np.testing.assert_array_equal(df["n_spikes"].to_numpy(),
                                  res.counts[df.index.to_numpy()])
    np.testing.assert_allclose(df["rate_hz"].to_numpy(),
                               res.counts[df.index.to_numpy()] / (0.25 * 2))
    z = np.load(tmp_path / "bins.npz", allow_pickle=False)
    assert set(z.files) == {"neuron_idx", "bins", "bin_ms", "t0_ms"}
    assert z["bins"].dtype == np.int16 and z["neuron_idx"].dtype == np.int32
    meta = json.loads((tmp_path / "meta.json").read_text(encoding="utf-8"))
    assert meta["n_neurons"] == 300 and meta["total_spikes"] == res.counts.sum()


def test_load_round_trips(tmp_path):
    res = _result(seed=3)
    S.save(res, tmp_path, min_rate=0.0)  # every neuron kept
    back = S.load(tmp_path)
    np.testing.assert_array_equal(back.counts, res.counts)
    np.testing.assert_array_equal(back.bins, res.bins)
    assert back.config.t_run_s == 0.25 and back.config.n_trials == 2
    assert back.stimuli[0].name == "a" and back.stimuli[0].rate_hz == 120.0
    np.testing.assert_array_equal(back.silenced, [3, 4])


def test_rate_table_carries_neuron_table_columns():
    res = _result(n=50)
    cols = list(S.Result.NEURON_COLUMNS)
    res.neurons = pd.DataFrame({c: ["%s%d" % (c, i) for i in range(50)]
                                for c in cols})
    df = res.rate_table(0.5)
    assert list(df.columns) == cols + ["rate_hz", "n_spikes"]
    assert (df["root_id"] == ["root_id%d" % i for i in df.index]).all()


def test_region_timeline_rates():
    """Mean rate per region and bin (sim.py:294-306), from a neuron table
    with a region column: neurons 0, 1 in 'AL', 2 in 'MB', 3 in 'AL';
    bins of 10 ms, 2 trials.  Bin 0 of AL holds 2 + 4 + 0 spikes over 3
    neurons: 6 / (3 * 0.01 s * 2) = 100 Hz."""
    bins = np.array([[2, 4, 1, 0], [0, 0, 3, 6]], np.int16)
    cfg = S.RunConfig(t_run_s=0.02, bin_ms=10.0, n_trials=2)
    nt = pd.DataFrame({"region": ["AL", "AL", "MB", "AL"]})
    res = S.Result(bins.sum(axis=0).astype(np.int64), bins, cfg, [],
                   np.zeros(0, np.int64), nt, {})
    tl = res.region_timeline()
    assert list(tl.index) == ["AL", "MB"] and tl.shape == (2, 2)
    np.testing.assert_allclose(tl.loc["AL"].to_numpy(), [100.0, 100.0])
    np.testing.assert_allclose(tl.loc["MB"].to_numpy(), [50.0, 150.0])
    # without a table, the connectome's neuropil groups name the regions
    res2 = S.Result(res.counts, bins, cfg, [], np.zeros(0, np.int64),
                    np.arange(4), {})
    res2.groups = np.array([1, 1, 0, 1])
    tl2 = res2.region_timeline()
    np.testing.assert_allclose(tl2.loc["group_1"].to_numpy(), [100.0, 100.0])


def test_state_and_tape_files_round_trip(tmp_path):
    """Snapshot, tape and trial-state files are plain .npz (no pickles)
    and read back exactly."""
    import torch

    from flybrain_amd import device as D

    state = {"v": np.linspace(-60, -40, 12, dtype=np.float32).reshape(2, 6),
             "ptr": np.int64(5), "gen": np.arange(16, dtype=np.uint8),
             "calls": np.int64(105)}
    S.save_state(state, tmp_path / "s.npz")
    back = S.load_state(tmp_path / "s.npz")
    assert set(back) == set(state)
    for k in state:
        np.testing.assert_array_equal(back[k], state[k])
    assert int(back["ptr"]) == 5

    steps = np.array([4, 0, 4, 9]), np.array([7, 1, 2, 7]), \
        np.array([1, 0, 1, 0])
    D.save_tape(tmp_path / "t.npz", *steps, n=10, batch=2,
                seeds=np.array([3, 4], np.uint64))
    z = D.load_tape(tmp_path / "t.npz")
    assert (z["n_steps"], z["batch"], z["n"]) == (10, 2, 10)
    got = sorted(zip(z["step"], z["neuron"], z["trial"]))
    assert got == sorted(zip(*steps))
    np.testing.assert_array_equal(z["seed"], [3, 4])
    np.load(tmp_path / "t.npz", allow_pickle=False)  # nothing pickled
    with pytest.raises(IndexError):
        D.save_tape(tmp_path / "bad.npz", [0], [10], [0], n=10, batch=2)
    with pytest.raises(ValueError):
        D.load_tape(tmp_path / "s.npz")

    ts = {"v": torch.arange(5, dtype=torch.float32), "g": torch.zeros(5),
          "st": torch.full((5,), 22, dtype=torch.uint8),
          "counts": torch.arange(5, dtype=torch.int32),
          "inflight": torch.ones(20, 5, dtype=torch.int32),
          "fl": np.zeros(5, np.uint8), "vhold": np.zeros(5, np.float32),
          "steps_done": 40, "steps_left": 60, "seed": 2**63 + 5,
          "trial_id": 0, "n": 5}
    D.save_trial_state(tmp_path / "x.npz", ts)
    tb = D.load_trial_state(tmp_path / "x.npz")
    for k, v in ts.items():
        if isinstance(v, torch.Tensor):
            assert torch.equal(tb[k], v), k
        elif isinstance(v, np.ndarray):
            np.testing.assert_array_equal(tb[k], v)
        else:
            assert tb[k] == v, k


def test_default_connectome_follows_the_reference_locations(tmp_path,
                                                             monkeypatch):
    """Without weights, the reference's cache/weights_coo.pt (under
    $FLYBRAIN_ROOT) is loaded, safely (weights_only); with nothing there
    the error names the paths."""
    import torch

    from flybrain_amd import connectome as C

    monkeypatch.setenv("FLYBRAIN_ROOT", str(tmp_path))
    with pytest.raises(FileNotFoundError, match="weights_coo.pt"):
        C.load_default()
    conn = C.synthetic(300, None, mean_out=10, n_groups=0, seed=4)
    (tmp_path / "cache").mkdir()
    torch.save(C.to_reference_coo(conn), tmp_path / "cache" / "weights_coo.pt")
    got = C.load_default()
    assert got.n == conn.n and got.fingerprint() == conn.fingerprint()
