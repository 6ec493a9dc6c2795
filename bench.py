"""Benchmark: simulated trial-seconds per wall-second on one MI355X.

Workloads (138,639 neurons, 15,091,983 edges, synthetic, seeded):
  headline    sensory_b256: B = 256 trials, four sensory index sets driven
              at piecewise-constant U(0, 300) Hz rates (10 ms segments),
              per-trial m_out masks, per-trial Philox seeds
  b1          plain_b1: B = 1, 20 neurons at 150 Hz -- the plain engine
  plain_b256  the plain_b1 drive replicated over B = 256 trials with
              distinct seeds: the like-for-like B = 256 : B = 1 ratio
              (every trial is the plain workload; batched = alone is a GPU
              test)
  b64         trials_b64: B = 64 trials, each driving its own 20-200
              neurons at U(50, 250) Hz with its own Philox seed

Every leg first runs ``--preroll`` untimed steps from rest (working
activity; independent of ``--warmup``), then ``--warmup`` untimed steps,
then the timed steps.  Windows are 20 steps; a HIP graph of
min(``--graph-windows``, steps // 20) windows is captured, so the timed
steps are graph replays (plus eager windows for a remainder).

A bench "step" is one 0.1 ms model step of every trial.  ``value`` is
B * steps * 1e-4 / wall seconds of the timed region (inputs and state
resident in HBM).  ``roofline`` is the dominant kernel's: ``achieved`` is
the per-window byte model (``window_bytes``: what one launch must move)
over the launch's average duration from HIP events on its stream;
``traffic`` is the HBM bytes per launch rocprofv3 counted (FETCH_SIZE and
WRITE_SIZE, committed under profiles/); ``valu_issue_frac`` is the
profiled VALU instructions per launch over what the chip issues in that
time.  ``b1.roofline`` puts the plain B = 1 kernel against the same peak,
by its per-window byte model and by the scope table's per-step formula.  ``cpu_baseline`` times the oracle's torch restatement of the
reference step (csrmv via torch.sparse.mm) on this host's cores.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (ROOT / "fly-in-silico_amd", ROOT / "oracle"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

HBM_PEAK_GBS = 8000.0
# wave64 VALU instructions the chip can issue per second: 256 CUs x 4 SIMDs
# x 2.4 GHz / 2 cycles per wave64 instruction (the 157.3 TFLOP/s fp32 peak
# as 2-flop FMAs over 64 lanes)
VALU_ISSUE_PER_S = 1.2288e12
# the scope table's dense term: bytes per neuron-trial-step of the update
# priced with a state round trip every step (reported as effective_gbs)
BETA = 26
# instructions of the plain recurrence per neuron-trial-step (gfx950
# assembly of run_window_plain)
PLAIN_STEP_VALU = 18


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed(fn, sync):
    sync()
    t0 = time.perf_counter()
    fn()
    sync()
    return time.perf_counter() - t0


def stages_of(runner):
    """(name, launch) of one window of ``runner``'s road: the drive launch
    (unless fused) and push update + scatter (one fused launch below 64
    trials) or the pull road's fused window kernel."""
    s = runner.sim
    W = runner.window
    st = [] if runner.fused else [
        ("drive", lambda: runner.table is not None
         and s.drive_table(runner.table, W))]
    if s.road == "pull":
        st.append(("pull", lambda: s.pull(W)))
    elif s.bstride < 64:
        st.append(("update_scatter", lambda: s.update_scatter(W)))
    else:
        st += [("update", lambda: s.update(W)),
               ("scatter", lambda: s.scatter())]
    return st


def kernel_times(runner, windows):
    """Median durations (ms) of each launch of a window over ``windows``
    eager windows, bracketed by HIP events on the launch stream (the
    median: a window delayed by the host now and then must not stand for
    the launch)."""
    import torch

    stages = stages_of(runner)
    ev = {k: [] for k, _ in stages}
    stream = torch.cuda.current_stream()
    # untimed windows first: after the host-side pause that follows the
    # timed region the first eager windows run slow while the GPU ramps
    # back up
    for _ in range(windows // 2):
        for _, fn in stages:
            fn()
    for _ in range(windows):
        marks = [torch.cuda.Event(enable_timing=True)
                 for _ in range(len(stages) + 1)]
        marks[0].record(stream)
        for i, (k, fn) in enumerate(stages):
            fn()
            marks[i + 1].record(stream)
            ev[k].append((marks[i], marks[i + 1]))
    torch.cuda.synchronize()
    out = {}
    for k, v in ev.items():
        d = sorted(a.elapsed_time(b) for a, b in v)
        out[k] = d[len(d) // 2]
    return out


def popcount64(a):
    import numpy as np
    return np.bitwise_count(np.ascontiguousarray(a).view(np.uint64))


def pull_window_bytes(sim, conn, W):
    """Bytes one pull-road window of ``W`` steps starting now must move
    (what the kernel reads and writes once per window; quiet neuron-trials
    are priced like the others):

      state       23 B per neuron-trial: v, g read + written (16), the
                  refractory byte read + written (2), flags (1), the
                  window's kick bits (4)
      in-edges    per tile part: 4 B source + 2 B count + 8 B act word per
                  in-edge, 4 B offset per neuron
      records     per active (in-edge, horizon step) pair -- a step at
                  which the source sent in some trial of the part -- the
                  part's 16-byte trial list (parts of 128 and 256 trials),
                  plus 8 B per 64-trial word of the part where the list
                  sends the reader to the words; 8 B per word otherwise
      own records W steps x bstride / 64 words x 8 B per neuron written,
                  and its act word read-modify-written (16 B per part)
      counts      8 B per neuron-trial that spiked in the window

    The active pairs are counted from the act words the window reads; the
    spiking neuron-trials from the records it writes (the window is run).
    """
    import numpy as np

    from flybrain_amd import _lib as L

    n, bs = sim.n, sim.bstride
    tile = sim.pull_tile
    tpn = min(bs, tile)
    parts = bs // tpn
    wpt = tpn // 64
    words = bs // 64
    t0 = sim.t
    hmask = 0
    for u in range(t0 - 20, t0 + W - 20):
        if u >= 0:
            hmask |= 1 << (u % L.REC_SLOTS)
    act = sim.act.cpu().numpy().reshape(n, parts).view(np.uint64)
    pairs_src = popcount64(act & np.uint64(hmask)).sum(axis=1)
    deg = conn.degree().astype(np.int64)
    pairs = int((deg * pairs_src).sum())
    lists = wpt >= 2 and tpn <= 256 and sim.rec_ids is not None
    over = 0  # pairs whose list sends the reader to the words
    if lists:
        cnt = sim.rec_ids.cpu().numpy().reshape(n, parts, L.REC_SLOTS, 16)
        cnt = cnt[:, :, :, 0]
        for u in range(t0 - 20, t0 + W - 20):
            if u < 0:
                continue
            j = u % L.REC_SLOTS
            valid = ((act >> np.uint64(j)) & np.uint64(1)).astype(bool)
            over += int((deg * (valid & (cnt[:, :, j] == 0xFF)).sum(axis=1)
                         ).sum())
    # run the window, then count the neuron-trials that sent in it (record
    # words of a step are valid where the part's act bit for it is set)
    sim.pull(W)
    rec = sim.rec.cpu().numpy().reshape(n, L.REC_SLOTS, words).view(np.uint64)
    act_after = sim.act.cpu().numpy().reshape(n, parts).view(np.uint64)
    sent = np.zeros((n, words), np.uint64)
    for u in range(t0, t0 + W):
        j = u % L.REC_SLOTS
        valid = ((act_after >> np.uint64(j)) & np.uint64(1)).astype(bool)
        valid_w = np.repeat(valid, words // parts, axis=1)
        sent |= np.where(valid_w, rec[:, j, :], np.uint64(0))
    spiking = int(popcount64(sent).sum())
    E = conn.n_edges
    parts_b = {
        "state": 23 * n * bs,
        "in_edges": parts * (14 * E + 4 * n),
        "records_read": (16 * pairs + 8 * wpt * over) if lists
                        else 8 * wpt * pairs,
        "records_written": 8 * W * words * n + 16 * n * parts,
        "counts": 8 * spiking,
    }
    return sum(parts_b.values()), parts_b, {"active_pairs": pairs,
                                            "pairs_read_as_words": over,
                                            "spiking_neuron_trials": spiking}


def push_window_bytes(sim, conn, W):
    """Bytes one push-road window (update + scatter, bstride < 64) must
    move: state 23 B per neuron-trial as on the pull road, the int32 ring
    slots the window reads (4 B per neuron-trial and step), and the
    scatter -- per spiking (neuron, window) row 8 B offsets + 6 B per
    edge, plus an 8 B int32 read-modify-write per edge, spiking trial and
    step it spiked at."""
    import numpy as np

    n, bs = sim.n, sim.bstride
    before = sim.counts.view(n, bs).cpu().numpy().copy()
    sim.step_launch(W)
    after = sim.counts.view(n, bs).cpu().numpy()
    spikes = (after.astype(np.int64) - before.astype(np.int64))[:, :sim.batch]
    deg = conn.degree().astype(np.int64)
    rows = spikes.sum(axis=1) > 0
    parts_b = {
        "state": 23 * n * bs,
        "ring_read": 4 * W * n * bs,
        "scatter_rows": int((8 + 6 * deg[rows]).sum()),
        "scatter_adds": int(8 * (deg[:, None] * spikes).sum()),
    }
    return sum(parts_b.values()), parts_b, {"spikes": int(spikes.sum())}


def window_model(runner, conn, windows=5):
    """Average bytes per window of the dominant launch, over ``windows``
    steady-state windows run here (eagerly, after the timed ones)."""
    s = runner.sim
    tot, parts_sum, extra_sum = 0, {}, {}
    for _ in range(windows):
        if runner.table is not None and not runner.fused:
            s.drive_table(runner.table, runner.window)
        if s.road == "pull":
            b, parts, extra = pull_window_bytes(s, conn, runner.window)
        else:
            b, parts, extra = push_window_bytes(s, conn, runner.window)
        tot += b
        for k, v in parts.items():
            parts_sum[k] = parts_sum.get(k, 0) + v
        for k, v in extra.items():
            extra_sum[k] = extra_sum.get(k, 0) + v
    return (tot // windows, {k: v // windows for k, v in parts_sum.items()},
            {k: v / windows for k, v in extra_sum.items()})


def scope_table_bytes(conn, n, B, W, spikes_per_trial_step):
    """The scope table's per-step formula for a window of W steps:
    N*B*26 + rows + edge atomics, with the window's average spikes spread
    over average-degree rows -- the figure behind ``effective_gbs``."""
    deg = float(conn.degree().mean())
    spikes = spikes_per_trial_step * B * W
    return int(BETA * n * B * W + spikes * (8 + 6 * deg) + 8 * spikes * deg)


def serve_bench(net, conn, args, rate, count):
    """Serving workload: Poisson arrivals of simulation requests
    (500-20,000 steps, 20-200 driven neurons at 50-250 Hz, 0-50 silenced,
    own seed) served by continuous batching in resident slots, and the same
    requests run back to back by the plain B = 1 engine.  Both in real
    time; results must be identical request by request."""
    import numpy as np

    from flybrain_amd.serve import Server, SimRequest, poisson_requests

    reqs = poisson_requests(conn.n, count, rate, seed=1)

    def copies():
        return [SimRequest(**vars(r)) for r in reqs]

    warm = [SimRequest(rid=-1 - k, steps=200, drive=reqs[k].drive,
                       rate_hz=reqs[k].rate_hz, seed=k) for k in range(2)]
    out = {"workload": "poisson_requests", "rate_req_s": rate,
           "requests": len(reqs),
           "mean_steps": float(np.mean([r.steps for r in reqs]))}
    runs = {}
    for name, kw in (("batched", dict(slots=args.serve_slots,
                                      shard=args.serve_shard)),
                     ("b1_back_to_back", dict(slots=1, road="push"))):
        srv = Server(net, max_drive=200, graph_windows=args.graph_windows,
                     **kw)
        srv.serve(warm, realtime=False)
        res, st = srv.serve(copies(), realtime=True)
        runs[name] = {r.rid: r.counts for r in res}
        st["road"] = srv.road
        st["shard"] = srv.shard
        out[name] = st
        log("serve %s @ %g req/s: %s" % (name, rate, st))
        del srv
    a, b = runs["batched"], runs["b1_back_to_back"]
    out["results_identical"] = bool(
        a.keys() == b.keys() and all(np.array_equal(a[k], b[k]) for k in a))
    return out


def _profile_rows(pattern, kernel, counters):
    """Mean value per dispatch of ``counters`` for ``kernel`` at its largest
    grid, from the rocprofv3 summaries committed under profiles/."""
    import glob

    best = {}
    for path in sorted(glob.glob(str(ROOT / "profiles" / pattern))):
        for line in open(path):
            parts = line.split()
            if len(parts) < 5 or parts[-3] not in counters:
                continue
            head = line[:60]
            if (kernel + "(" not in head and kernel + "<" not in head
                    and not (kernel.endswith(">") and kernel in head)):
                continue
            grid, name, val = int(parts[-4]), parts[-3], float(parts[-1])
            if grid >= best.get(name, (0, 0))[0]:
                best[name] = (grid, val)
    return {k: v for k, (_, v) in best.items()}


def profiled_traffic(kernel):
    """HBM bytes per launch of ``kernel`` from the FETCH_SIZE / WRITE_SIZE
    summaries (KiB per dispatch, scaled by what the same counters read for
    a plain 256 MiB copy: FETCH x 2.0, WRITE x 1.0).  None when not
    profiled."""
    got = _profile_rows("*_hbm_counters.md", kernel,
                        ("FETCH_SIZE", "WRITE_SIZE"))
    if "FETCH_SIZE" not in got or "WRITE_SIZE" not in got:
        return None
    return int(1024 * (2.0 * got["FETCH_SIZE"] + 1.0 * got["WRITE_SIZE"]))


def profiled_valu(kernel):
    """SQ_INSTS_VALU per launch of ``kernel`` from the committed SQ
    counter summaries, or None."""
    return _profile_rows("*_sq_counters.md", kernel,
                         ("SQ_INSTS_VALU",)).get("SQ_INSTS_VALU")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg(conn, wl, seconds, max_steps=None):
    """The oracle's torch restatement of the reference step on this host,
    B = 1, driven like ``wl`` (Bernoulli outcomes drawn on the host)."""
    import numpy as np
    import torch

    import ref_torch

    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    eng = ref_torch.TorchRef(conn, 1, no_refractory=wl.drive_neuron)
    rng = np.random.default_rng(0)
    p = float(np.float32(wl.rate_hz[0, 0]) * np.float32(1e-4))
    bits = np.zeros((1, conn.n), np.uint8)
    for _ in range(5):
        eng.step(bits)
    steps = 0
    t0 = time.perf_counter()
    while (time.perf_counter() - t0 < seconds
           and (max_steps is None or steps < max_steps)):
        bits[:] = 0
        kick = wl.drive_neuron[rng.random(wl.drive_neuron.size) < p]
        bits[0, kick] = 1
        eng.step(bits)
        steps += 1
    wall = time.perf_counter() - t0
    return steps, wall, cores


def cpu_baseline(conn, seconds):
    """Configs 2 (138,639 neurons; a bounded sample of the B = 1 workload)
    and 1 (10,000 neurons, 1,000 steps) on the host's cores."""
    from flybrain_amd import connectome as C
    from flybrain_amd import workloads as WL

    steps, wall, cores = cpu_leg(conn, WL.plain_b1(conn.n, seed=0), seconds)
    small = C.small()
    s1, w1, _ = cpu_leg(small, WL.plain_b1(small.n, seed=0), 120.0,
                        max_steps=1000)
    return {"value": round(steps * 1e-4 / wall, 5), "unit": "trial-s/s",
            "cores": cores, "kind": "port",
            "sample": "%d steps of the B=1 workload (138,639 neurons, 20 "
                      "neurons at 150 Hz), torch CPU csrmv restatement of "
                      "the reference step, %.1f s" % (steps, wall),
            "ms_per_step": round(1e3 * wall / steps, 3),
            "cpu_model": cpu_model(),
            "config1": {"workload": "10,000 neurons, %d edges, 20 neurons "
                                    "at 150 Hz, B=1" % small.n_edges,
                        "steps": s1, "wall_s": round(w1, 3),
                        "value": round(s1 * 1e-4 / w1, 5),
                        "ms_per_step": round(1e3 * w1 / s1, 4)}}


def summarize(paths):
    """Per-kernel summary of rocprofv3 databases (kernel trace or --pmc).

    Dispatches are grouped by kernel name and grid size (which tells the
    workloads apart).  For counter databases the mean value per dispatch
    is printed.
    """
    import sqlite3

    for path in paths:
        con = sqlite3.connect(path)
        print("# %s" % path)
        rows = con.execute(
            "select name, grid_x, count(*), avg(duration), min(duration), "
            "max(duration), sum(duration) from kernels group by name, grid_x "
            "order by sum(duration) desc").fetchall()
        if rows and rows[0][2]:
            print("%-60s %10s %7s %12s %12s %12s %12s" % (
                "kernel", "grid", "calls", "avg_us", "median_us", "min_us",
                "max_us"))
            for name, gx, n, avg, mn, mx, tot in rows[:14]:
                durs = sorted(d for (d,) in con.execute(
                    "select duration from kernels where name = ? and "
                    "grid_x = ?", (name, gx)))
                med = durs[len(durs) // 2] if durs else 0.0
                print("%-60s %10d %7d %12.3f %12.3f %12.3f %12.3f" % (
                    _short(name), gx, n, avg / 1e3, med / 1e3, mn / 1e3,
                    mx / 1e3))
        pmc = con.execute(
            "select kernel_name, grid_size, counter_name, count(*), avg(value) "
            "from counters_collection group by kernel_name, grid_size, "
            "counter_name order by kernel_name").fetchall()
        if pmc:
            print("%-60s %12s %-12s %7s %16s" % ("kernel", "grid", "counter",
                                                 "calls", "mean_value"))
            for name, gs, cn, n, val in pmc:
                if n and ("fis" in name or "k_" in name
                          or "elementwise" in name):
                    print("%-60s %12d %-12s %7d %16.1f" % (
                        _short(name), gs, cn, n, val))
        print()


def b1_roofline(rr, conn):
    """The plain B = 1 kernel against the HBM roofline, two ways: the bytes
    one window must move (``push_window_bytes``: state once per window, the
    ring slots it reads, the scatter) over the launch's median time, and
    the scope table's per-step formula (26 B per neuron-step priced at every
    step, plus the scatter) over the measured time per step -- the
    definition the plain kernel's 40 % target is stated in."""
    r = rr["runner"]
    W = r.window
    launch = rr["kt"]["update_scatter"]
    wb, parts, extra = window_model(r, conn)
    spts = rr["spikes"] / (rr["wl"].batch * rr["steps"])
    st = scope_table_bytes(conn, conn.n, 1, W, spts) / W  # per step
    step_s = rr["wall"] / rr["steps"]
    achieved = wb / (launch * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": "k_update<20, true>",
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "launch_ms": round(launch, 5), "window_bytes": int(wb),
            "window_bytes_parts": parts,
            "achieved": round(achieved, 1),
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "scope_table_bytes_per_step": int(st),
            "scope_table_gbs": round(st / step_s / 1e9, 1),
            "scope_table_frac": round(st / step_s / 1e9 / HBM_PEAK_GBS, 4)}


def _short(name):
    name = name.replace("(anonymous namespace)::", "")
    return name[:60]


def run_leg(name, net, conn, args, wl, steps):
    """Preroll, warm-up and ``steps`` timed steps of ``wl``; kernel times
    from HIP events afterwards."""
    from flybrain_amd.batch import BatchRunner

    tuned = None
    if args.road == "tuned":
        from flybrain_amd.autotune import tune
        t1 = time.perf_counter()
        tuned = tune(net, wl)
        log("%s: tuned %s in %.1f s, ms per window %s"
            % (name, tuned["best"], time.perf_counter() - t1,
               tuned["times"]))
    r = BatchRunner(net, wl, counts=True, road=args.road)
    W = r.window
    # a graph that fits the timed steps, so that they replay it
    gw = min(args.graph_windows, steps // W) if not args.no_graph else 0
    if gw > 0:
        r.capture(gw)
    use = gw > 0
    import torch
    sync = torch.cuda.synchronize
    r.advance(args.preroll, use_graph=use)
    r.advance(args.warmup, use_graph=use)
    # spike totals summed on the device: a host-side sum of the 35 M
    # counts leaves the GPU idle for tens of milliseconds right before the
    # timed steps, and it starts them at lowered clocks
    total_spikes = lambda: int(r.sim.counts.sum(dtype=torch.int64))  # noqa: E731
    c0 = total_spikes()
    wall = timed(lambda: r.advance(steps, use_graph=use), sync)
    spikes = total_spikes() - c0
    kt = kernel_times(r, 20)
    res = dict(wl=wl, runner=r, wall=wall, steps=steps, spikes=spikes,
               kt=kt, tuned=tuned, graph_windows=gw,
               tps=wl.batch * steps * 1e-4 / wall)
    log("%s: B=%d %d steps %.4f s  %.3f us/step  %.2f trial-s/s  "
        "spikes/trial-step %.2f  kernels(ms) %s  graph %d windows"
        % (name, wl.batch, steps, wall, 1e6 * wall / steps, res["tps"],
           spikes / (wl.batch * steps),
           {k: round(v, 5) for k, v in kt.items()}, gw))
    return res


def leg_summary(rr):
    r = rr["runner"]
    return {"workload": rr["wl"].name, "batch": rr["wl"].batch,
            "value": round(rr["tps"], 3), "steps": rr["steps"],
            "ms_per_step": round(1e3 * rr["wall"] / rr["steps"], 6),
            "graph_windows": rr["graph_windows"],
            "road": r.sim.road,
            "pull_tile": r.sim.pull_tile if r.sim.road == "pull" else None,
            "gate": r.sim.gate,
            "autotuned": rr["tuned"]["times"] if rr["tuned"] else None,
            "spikes_per_trial_step": round(
                rr["spikes"] / (rr["wl"].batch * rr["steps"]), 3),
            "kernels_ms": {k: round(v, 5) for k, v in rr["kt"].items()}}


def main():
    if len(sys.argv) > 2 and sys.argv[1] == "--summarize":
        summarize(sys.argv[2:])
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--preroll", type=int, default=1000,
                    help="untimed steps from rest before the warm-up of "
                         "every leg (working activity)")
    ap.add_argument("--b1-steps", type=int, default=10_000)
    ap.add_argument("--b64-steps", type=int, default=10_000)
    ap.add_argument("--plain-steps", type=int, default=2_000,
                    help="timed steps of the plain_b256 leg")
    ap.add_argument("--graph-windows", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=("all", "b256", "b1", "b64",
                                       "plain_b256"),
                    default="all")
    ap.add_argument("--serve-requests", type=int, default=128,
                    help="requests of the 200 req/s serving run (0: skip "
                         "serving)")
    ap.add_argument("--serve-rate", type=float, default=200.0,
                    help="Poisson arrival rate of requests per second")
    ap.add_argument("--serve-low-rate", type=float, default=20.0)
    ap.add_argument("--serve-low-requests", type=int, default=48)
    ap.add_argument("--serve-slots", type=int, default=256)
    ap.add_argument("--serve-shard", type=int, default=16,
                    help="slots per shard (one DeviceSim each)")
    ap.add_argument("--road", choices=("tuned", "auto", "push", "pull"),
                    default="tuned",
                    help="tuned: the (road, pull workgroup shape) the "
                         "autotuner measures fastest for each workload "
                         "(cached); auto: pull road for 64..1024 trials, "
                         "push below")
    ap.add_argument("--no-graph", action="store_true",
                    help="eager launches (for per-dispatch counter runs)")
    ap.add_argument("--no-model", action="store_true",
                    help="skip the per-window byte model (profiling runs)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a 256 MiB elementwise copy (counter "
                         "calibration: known bytes read and written)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the whole batch runs on one "
              "MI355X (trials would shard across GPUs with no exchange).  "
              "--gpus %d is not run." % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import torch

    from flybrain_amd import connectome as C
    from flybrain_amd import workloads as WL
    from flybrain_amd.device import DeviceNet

    t0 = time.perf_counter()
    conn = C.full()
    log("connectome %d neurons, %d edges, %s (%.1f s)"
        % (conn.n, conn.n_edges, conn.fingerprint(), time.perf_counter() - t0))
    net = DeviceNet(conn, "cuda")
    out = {"metric": "sim trial-s per wall-s at 138k neurons, B=256 vs B=1; "
                     "p50/p99 req latency",
           "unit": "trial-s/s", "n_gpus": 1, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32+i32",
           "data": "synthetic"}

    total = lambda k: args.preroll + args.warmup + k  # noqa: E731
    makers = {
        "b256": (lambda: WL.sensory_b256(conn.n, seed=0,
                                         steps=total(args.steps)),
                 args.steps),
        "b1": (lambda: WL.plain_b1(conn.n, seed=0, steps=total(args.b1_steps)),
               args.b1_steps),
        "plain_b256": (lambda: WL.plain_replicated(
            conn.n, seed=0, batch=256, steps=total(args.plain_steps)),
            args.plain_steps),
        "b64": (lambda: WL.trials_b64(conn.n, seed=0,
                                      steps=total(args.b64_steps)),
                args.b64_steps),
    }
    names = (("b256", "b1", "plain_b256", "b64") if args.only == "all"
             else (args.only,))
    res = {}
    for name in names:
        mk, steps = makers[name]
        res[name] = run_leg(name, net, conn, args, mk(), steps)

    if args.calib_copy:
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        b = a + 1.0  # reads 256 MiB, writes 256 MiB
        torch.cuda.synchronize()
        log("calibration copy: 256 MiB read + 256 MiB written (%s)"
            % float(b[0]))
        del a, b

    head = res.get("b256") or res[names[0]]
    wl = head["wl"]
    r = head["runner"]
    n, B, W = conn.n, wl.batch, r.window
    kt = head["kt"]
    dom = max((k for k in kt if k != "drive"), key=lambda k: kt[k])
    hs = r.sim
    kname = {"update": "k_update", "update_scatter": "k_update",
             "pull": "k_pull<%d, %d, %d, %s>" % (W, hs.bstride, hs.pull_tile,
                                                 "true" if hs.gate else "false"),
             "scatter": "k_scatter_steps"}[dom]
    spts = head["spikes"] / (B * head["steps"])
    roof = {"bound": "hbm", "kernel": kname, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "launch_ms": round(kt[dom], 5)}
    if not args.no_model and dom in ("pull", "update_scatter"):
        wb, parts, extra = window_model(r, conn)
        achieved = wb / (kt[dom] * 1e-3) / 1e9
        roof.update({"achieved": round(achieved, 1),
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "window_bytes": int(wb), "window_bytes_parts": parts,
                     "window_model_counts": extra})
    traffic = profiled_traffic(kname)
    roof["traffic"] = traffic
    roof["traffic_gbs"] = (round(traffic / (kt[dom] * 1e-3) / 1e9, 1)
                           if traffic else None)
    valu = profiled_valu(kname)
    roof["valu_insts_per_launch"] = valu
    roof["valu_issue_frac"] = (round(valu / (VALU_ISSUE_PER_S
                                             * kt[dom] * 1e-3), 4)
                               if valu else None)
    st_b = scope_table_bytes(conn, n, B, W, spts)
    roof["effective_gbs"] = round(st_b / (kt[dom] * 1e-3) / 1e9, 1)
    roof["effective_bytes_note"] = (
        "scope-table formula, 26 B per neuron-trial-step priced every step "
        "plus the scatter term; not bytes the windowed kernel moves")
    floor_ms = n * B / 64 * PLAIN_STEP_VALU * W / VALU_ISSUE_PER_S * 1e3
    out.update({
        "value": round(head["tps"], 3),
        "steps": head["steps"], "warmup": args.warmup,
        "ms_per_step": round(1e3 * head["wall"] / head["steps"], 6),
        "config": {"workload": wl.name, "neurons": n,
                   "edges": conn.n_edges, "batch": B,
                   "road": hs.road,
                   "pull_tile": hs.pull_tile if hs.road == "pull" else None,
                   "gate": hs.gate,
                   "autotuned": (head["tuned"]["times"] if head["tuned"]
                                 else None),
                   "window_steps": W,
                   "graph_windows": head["graph_windows"],
                   "preroll_steps": args.preroll,
                   "connectome": conn.fingerprint()},
        "roofline": roof,
        "kernels_ms": {k: round(v, 5) for k, v in kt.items()},
        "kernel_ms_per_step": round(sum(kt.values()) / W, 6),
        "recurrence_floor_ms_per_window": round(floor_ms, 5),
    })
    for name, rr in res.items():
        if rr is not head:
            out[name] = leg_summary(rr)
    if "b1" in res and res["b1"] is not head and not args.no_model:
        out["b1"]["roofline"] = b1_roofline(res["b1"], conn)
    out["spikes_per_trial_step"] = {
        name: round(rr["spikes"] / (rr["wl"].batch * rr["steps"]), 3)
        for name, rr in res.items()}
    if "b1" in res:
        b1 = res["b1"]["tps"]
        if "b256" in res:
            out["ratio_b256_over_b1"] = round(res["b256"]["tps"] / b1, 3)
        if "plain_b256" in res:
            out["ratio_plain_b256_over_b1"] = round(
                res["plain_b256"]["tps"] / b1, 3)
    if args.serve_requests > 0 and args.only == "all":
        sv = serve_bench(net, conn, args, args.serve_rate,
                         args.serve_requests)
        out["serving"] = sv
        out["p50_ms"] = sv["batched"]["p50_ms"]
        out["p99_ms"] = sv["batched"]["p99_ms"]
        if args.serve_low_requests > 0:
            out["serving_low_rate"] = serve_bench(
                net, conn, args, args.serve_low_rate,
                args.serve_low_requests)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(conn, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
