"""Connectome generator, formats, silencing and exactness (CPU)."""

from __future__ import annotations

import numpy as np
import pytest

from flybrain_amd import connectome as C


@pytest.fixture(scope="module")
def mid():
    # same laws as the full graph at 1/7 of its size
    return C.synthetic(20_000, 2_180_000, n_groups=36, seed=5)


def test_sizes_and_structure(mid):
    c = mid
    assert c.n_edges == 2_180_000
    pre = np.repeat(np.arange(c.n), c.degree())
    assert not np.any(pre == c.cols), "self edge"
    key = pre * c.n + c.cols
    assert np.unique(key).size == key.size, "duplicate edge"
    assert np.all(np.diff(c.cols.astype(np.int64))[np.diff(pre) == 0] > 0)
    assert c.degree().min() >= 1


def test_degree_law(mid):
    d = mid.degree().astype(np.float64)
    assert abs(d.mean() - 109.0) < 0.5
    # log-normal sigma 1: std of log degree near 1 (rounding and the
    # rescale pull it slightly)
    assert 0.85 < np.log(d).std() < 1.1


def test_count_and_sign_laws(mid):
    v = mid.vals.astype(np.int64)
    a = np.abs(v)
    assert a.min() >= 1 and a.max() <= 2000
    # P(k) ~ k^-2 on [1, 2000]: P(1) = 1 / H(2000, 2) = 0.6084
    p1 = (a == 1).mean()
    assert abs(p1 - 0.6084) < 0.005
    assert abs(a.mean() - 4.97) < 0.3
    # sign is per presynaptic neuron
    pre = np.repeat(np.arange(mid.n), mid.degree())
    sign_of_pre = np.zeros(mid.n)
    sign_of_pre[pre] = np.sign(v)
    assert np.all(np.sign(v) == sign_of_pre[pre])
    assert abs((sign_of_pre > 0).mean() - 0.55) < 0.02


def test_within_group_fraction(mid):
    pre = np.repeat(np.arange(mid.n), mid.degree())
    same = mid.groups[pre] == mid.groups[mid.cols]
    assert abs(same.mean() - 0.8) < 0.03
    assert np.unique(mid.groups).size == 36


def test_seeded_and_stable():
    a = C.synthetic(3000, 200_000, n_groups=5, seed=9)
    b = C.synthetic(3000, 200_000, n_groups=5, seed=9)
    c = C.synthetic(3000, 200_000, n_groups=5, seed=10)
    assert a.fingerprint() == b.fingerprint() != c.fingerprint()


def test_small_workload_graph(golden):
    z = golden("small10k.npz")
    c = C.small()
    assert c.n == 10_000 and c.n_edges == int(z["n_edges"]) == 1_090_000
    assert c.fingerprint() == str(z["conn_fp"])


def test_reference_format_round_trip():
    c = C.synthetic(500, None, mean_out=20, n_groups=3, seed=1)
    obj = C.to_reference_coo(c)
    assert obj["indices"].shape == (2, c.n_edges)
    assert tuple(obj["shape"]) == (500, 500)
    back = C.from_reference_coo(obj)
    assert back.fingerprint() == c.fingerprint()
    # a torch sparse tensor, as load_weights returns, reads the same
    import torch

    w = torch.sparse_coo_tensor(obj["indices"], obj["values"], (500, 500))
    assert C.from_reference_coo(w).fingerprint() == c.fingerprint()
    dense = c.dense_counts()
    np.testing.assert_array_equal(w.to_dense().numpy(), dense.astype(np.float32))


def test_duplicates_are_summed_like_coalesce():
    c = C.from_edges(4, [0, 0, 1, 0], [1, 1, 2, 3], [2, 5, -3, 1])
    np.testing.assert_array_equal(c.offsets, [0, 2, 3, 3, 3])
    np.testing.assert_array_equal(c.cols, [1, 3, 2])
    np.testing.assert_array_equal(c.vals, [7, 1, -3])
    with pytest.raises(ValueError):
        C.from_edges(4, [0], [1], [0.5])


def test_silence_matches_dense_rule():
    c = C.synthetic(400, None, mean_out=30, n_groups=0, seed=2)
    sil = [3, 17, 200]
    w = c.dense_counts()
    w[sil, :] = 0
    w[:, sil] = 0
    np.testing.assert_array_equal(c.silence(sil).dense_counts(), w)
    assert c.silence([]) is c


def test_exactness_matches_reference(golden):
    """max row sum of |w| equals the reference's evspmv.exactness number."""
    z = golden("tiny_masks.npz")
    c = C.Connectome(int(z["n"]), z["conn_offsets"], z["conn_cols"],
                     z["conn_vals"])
    ex = c.exactness()
    assert ex["max_row_abs_sum"] == float(z["exactness_max_row"])
    assert ex["ok"]


def test_bounds_checked():
    with pytest.raises(ValueError):
        C.Connectome(3, np.array([0, 1, 2, 2]), np.array([1, 5], np.int32),
                     np.array([1, 1], np.int16))
    with pytest.raises(ValueError):
        C.Connectome(3, np.array([0, 2, 1, 2]), np.array([1, 2], np.int32),
                     np.array([1, 1], np.int16))


def test_save_load(tmp_path):
    c = C.synthetic(300, None, mean_out=10, n_groups=2, seed=3)
    C.save(c, tmp_path / "c.npz")
    d = C.load(tmp_path / "c.npz")
    assert d.fingerprint() == c.fingerprint()
    np.testing.assert_array_equal(d.groups, c.groups)


@pytest.mark.parametrize("bstride", [64, 128, 256, 1024])
@pytest.mark.parametrize("locality", [True, False])
def test_pull_layout_lists_every_in_edge_by_position(bstride, locality):
    """The pull road's neuron order is a permutation of aligned units of
    max(1, 256 / bstride) neurons (grouped by neuropil when asked), and the
    in-edge CSR by position holds exactly each neuron's row of W."""
    import torch

    from flybrain_amd.device import DeviceNet

    conn = C.synthetic(1003, None, mean_out=30, n_groups=5, seed=4)
    net = DeviceNet(conn, "cpu")
    net.locality = locality
    order, off, cols, vals = [t.numpy() for t in net.pull_layout(bstride)]
    assert net.pull_layout(bstride)[0] is not None  # cached
    q = max(1, 256 // bstride)
    assert np.array_equal(np.sort(order), np.arange(conn.n))
    full = conn.n // q
    units = order[:full * q].reshape(-1, q)
    assert (units[:, 0] % q == 0).all() and (np.diff(units, axis=1) == 1).all()
    g = conn.groups[order[:full * q:q]]
    if locality:
        assert (np.diff(g.astype(np.int64)) >= 0).all()
    else:
        assert np.array_equal(order, np.arange(conn.n))
    dense = conn.dense_counts()
    for k in range(0, conn.n, 17):
        got = np.zeros(conn.n, np.int64)
        np.add.at(got, cols[off[k]:off[k + 1]], vals[off[k]:off[k + 1]])
        assert np.array_equal(got, dense[order[k]]), k
    assert torch.is_tensor(net.pull_layout(bstride)[1])


def test_flywire_tables_round_trip(tmp_path):
    """The reference's FlyWire layout (connectivity parquet with the index
    and 'Excitatory x Connectivity' columns, completeness CSV giving N)
    round-trips, and duplicate rows are summed as coalesce() sums them."""
    import pandas as pd

    c = C.synthetic(400, None, mean_out=20, n_groups=0, seed=4)
    C.to_flywire_tables(c, tmp_path / "conn.parquet", tmp_path / "comp.csv")
    d = C.from_flywire_tables(tmp_path / "conn.parquet", tmp_path / "comp.csv")
    assert d.n == c.n and d.fingerprint() == c.fingerprint()
    df = pd.read_parquet(tmp_path / "conn.parquet")
    assert list(df.columns) == list(C.FLYWIRE_COLUMNS)
    # a duplicated connection row: its counts add up
    dup = pd.concat([df, df.iloc[:5]], ignore_index=True)
    dup.to_parquet(tmp_path / "dup.parquet", index=False)
    e = C.from_flywire_tables(tmp_path / "dup.parquet", n=c.n)
    pre = np.repeat(np.arange(c.n), c.degree())
    for k in range(5):
        i, j = int(df.iloc[k, 0]), int(df.iloc[k, 1])
        row = slice(e.offsets[i], e.offsets[i + 1])
        got = e.vals[row][e.cols[row] == j]
        assert got.tolist() == [2 * int(df.iloc[k, 2])]
    assert e.n_edges == c.n_edges and pre.size == c.n_edges
