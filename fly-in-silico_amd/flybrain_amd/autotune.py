"""Trial-block shape picked by measurement.

A batch of trials runs on one of two roads (push: event scatter into an
int32 ring; pull: a fused gather + update per window, 64..1024 trials) and,
on the pull road, with one of the compiled workgroup shapes
(``_lib.PULL_TILES`` neuron-trials per workgroup) and with or without input
gating (tiles no input reaches skip their in-edge walk, at the price of
stamping every sender's targets).  Which is fastest
depends on the batch, the connectome and how active it is, so
:func:`tune` times every candidate on the workload itself -- a short
warm-up to reach its working activity, then a few eager windows bracketed
by device events -- and keeps the fastest.  Results are cached in a JSON
file (``$FLYBRAIN_TUNE_CACHE``, default ``~/.cache/flybrain_amd/tune.json``)
keyed by device, the library build (ABI version and a hash of libfis.so),
connectome fingerprint, workload name, batch and a hash of the workload's
arrays; the cache is only an optimisation (an unwritable location is
skipped).
"""

from __future__ import annotations

import json
import os
from pathlib import Path

from . import _lib as L
from .device import PULL_MIN_BATCH


def cache_path():
    p = os.environ.get("FLYBRAIN_TUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "flybrain_amd" / "tune.json"


def candidates(batch):
    """(road, pull_tile, gate) choices that can run ``batch`` trials: the
    push road, and the pull road with each compiled workgroup shape,
    with and without input gating."""
    out = [("push", None, False)]
    if PULL_MIN_BATCH <= batch <= L.PULL_MAX:
        out += [("pull", t, g) for g in (False, True) for t in L.PULL_TILES]
    return out


def _label(road, tile, gate):
    return "%s/%s%s" % (road, tile or "-", "/gated" if gate else "")


def library_fingerprint():
    """ABI version and a hash of the built libfis.so: a choice measured with
    other kernels is not reused after a rebuild."""
    import hashlib

    h = hashlib.sha256()
    try:
        h.update(L.LIB_PATH.read_bytes())
    except OSError:
        h.update(b"missing")
    return "abi%d-%s" % (L.ABI_VERSION, h.hexdigest()[:16])


def workload_fingerprint(wl):
    """Hash of everything that sets a workload's activity: seeds, drive
    entries, rate schedule and m_out pairs (not only its name)."""
    import hashlib

    import numpy as np

    h = hashlib.sha256()
    h.update(("%s|%d|%d" % (wl.name, wl.batch, wl.seg_steps)).encode())
    for a in (wl.seeds, wl.drive_neuron, wl.drive_trial, wl.drive_group,
              wl.rate_hz, wl.out_neuron, wl.out_trial):
        a = np.ascontiguousarray(a)
        h.update(str((a.dtype.str, a.shape)).encode())
        h.update(a.tobytes())
    return h.hexdigest()[:16]


def cache_key(net, wl):
    import torch
    dev = torch.cuda.get_device_name(net.device) if net.device.type == "cuda" \
        else str(net.device)
    return "%s|%s|%s|%s|%d|%s" % (dev, library_fingerprint(),
                                  net.conn.fingerprint(), wl.name, wl.batch,
                                  workload_fingerprint(wl))


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, table):
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(".tmp")
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def time_candidate(net, wl, road, pull_tile, gate=False, warmup_steps=200,
                   windows=10):
    """Mean wall time of one 20-step window (ms) of ``wl`` on (road,
    pull_tile, gate), eager launches, after ``warmup_steps`` steps."""
    import torch

    from .batch import BatchRunner

    r = BatchRunner(net, wl, counts=False, road=road, pull_tile=pull_tile,
                    gate=gate)
    r.advance(warmup_steps, use_graph=False)
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    r.advance(windows * r.window, use_graph=False)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / windows
    del r
    return ms


def tune(net, wl, cache=True, timer=None, **kw):
    """Pick the fastest (road, pull_tile) for ``wl`` on ``net``.

    Returns ``{"best": {"road", "pull_tile", "gate", "ms_per_window"},
    "times": {"road/tile[/gated]": ms}}``.  ``timer(net, wl, road, tile,
    gate) -> ms`` replaces the device measurement (tests); ``kw`` goes to
    :func:`time_candidate`.
    """
    path = cache_path()
    key = cache_key(net, wl)
    table = _load(path) if cache else {}
    if key in table:
        return table[key]
    timer = timer or (lambda n, w, road, tile, gate: time_candidate(
        n, w, road, tile, gate, **kw))
    times = {}
    best = None
    for road, tile, gate in candidates(wl.batch):
        ms = float(timer(net, wl, road, tile, gate))
        times[_label(road, tile, gate)] = round(ms, 4)
        if best is None or ms < best["ms_per_window"]:
            best = {"road": road, "pull_tile": tile, "gate": gate,
                    "ms_per_window": ms}
    result = {"best": best, "times": times}
    if cache:
        table[key] = result
        _store(path, table)
    return result
