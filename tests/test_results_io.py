"""Run results on disk (CPU): save() writes the reference's files --
meta.json, rates.parquet (index = neuron, rate_hz, n_spikes), bins.npz --
and load() rebuilds the Result (lab/flybrain/sim.py:1120-1174)."""

from __future__ import annotations

import json

import numpy as np
import pandas as pd

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
