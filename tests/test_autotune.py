"""The trial-block autotuner's host logic: candidates, choice, cache."""

from __future__ import annotations

import json
from types import SimpleNamespace

import torch

from flybrain_amd import _lib as L
from flybrain_amd import autotune as T


def fake(batch, fp="c0ffee", seed=0):
    from flybrain_amd import workloads as WL

    net = SimpleNamespace(device=torch.device("cpu"),
                          conn=SimpleNamespace(fingerprint=lambda: fp))
    wl = WL.trials_b64(500, seed=seed, batch=batch, steps=100)
    return net, wl


def test_candidates_follow_the_roads():
    assert T.candidates(1) == [("push", None, False)]
    assert T.candidates(63) == [("push", None, False)]
    got = T.candidates(256)
    assert got[0] == ("push", None, False)
    pull = [(t, g) for r, t, g in got if r == "pull"]
    assert sorted(pull) == sorted((t, g) for t in L.PULL_TILES
                                  for g in (False, True))
    assert T.candidates(L.PULL_MAX + 1) == [("push", None, False)]


def test_tune_picks_the_fastest_and_caches(tmp_path, monkeypatch):
    monkeypatch.setenv("FLYBRAIN_TUNE_CACHE", str(tmp_path / "t.json"))
    net, wl = fake(256)
    cost = {("push", None, False): 8.0, ("pull", 128, False): 1.2,
            ("pull", 256, False): 1.1, ("pull", 512, False): 1.4,
            ("pull", 128, True): 1.3, ("pull", 256, True): 0.9,
            ("pull", 512, True): 1.5}
    calls = []

    def timer(n, w, road, tile, gate):
        calls.append((road, tile, gate))
        return cost[(road, tile, gate)]

    res = T.tune(net, wl, timer=timer)
    assert res["best"]["road"] == "pull" and res["best"]["pull_tile"] == 256
    assert res["best"]["gate"] is True
    assert res["times"]["push/-"] == 8.0 and len(res["times"]) == 7
    assert res["times"]["pull/256/gated"] == 0.9
    assert sorted(calls, key=str) == sorted(cost, key=str)
    # cached: no second measurement, same answer
    again = T.tune(net, wl, timer=lambda *a: 1 / 0)
    assert again == json.loads(json.dumps(res))
    table = json.loads((tmp_path / "t.json").read_text())
    assert list(table) == [T.cache_key(net, wl)]
    # another connectome or batch is measured anew
    net2, _ = fake(256, fp="beef")
    assert T.tune(net2, wl, timer=lambda n, w, r, t, g: 2.0 if r == "push"
                  else 3.0)["best"]["road"] == "push"


def test_unwritable_cache_is_skipped(tmp_path, monkeypatch):
    blocker = tmp_path / "file"
    blocker.write_text("x")
    monkeypatch.setenv("FLYBRAIN_TUNE_CACHE", str(blocker / "sub" / "t.json"))
    net, wl = fake(32)
    res = T.tune(net, wl, timer=lambda *a: 5.0)
    assert res["best"] == {"road": "push", "pull_tile": None, "gate": False,
                           "ms_per_window": 5.0}


def test_cache_key_covers_library_and_workload_content():
    """A choice is not reused for another build of the library or for a
    workload of the same name and batch with other inputs."""
    net, wl = fake(64)
    _, other = fake(64, seed=1)
    assert wl.name == other.name and wl.batch == other.batch
    assert T.cache_key(net, wl) != T.cache_key(net, other)
    assert T.cache_key(net, wl) == T.cache_key(net, fake(64)[1])
    assert T.library_fingerprint() in T.cache_key(net, wl)
    assert T.library_fingerprint().startswith("abi%d-" % L.ABI_VERSION)
