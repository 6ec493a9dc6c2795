"""Benchmark: simulated trial-seconds per wall-second on one MI355X.

Workloads (138,639 neurons, 15,091,983 edges, synthetic, seeded):
  headline  sensory_b256: B = 256 trials, four sensory index sets driven at
            piecewise-constant U(0, 300) Hz rates (10 ms segments), per-trial
            m_out masks, per-trial Philox seeds; windows of 20 steps replayed
            from one captured HIP graph
  b1        plain_b1: B = 1, 20 neurons at 150 Hz -- the plain engine
  b64       trials_b64: B = 64 trials, each driving its own 20-200 neurons
            at U(50, 250) Hz with its own Philox seed; road and pull-road
            workgroup shape (trial-block width) autotuned

A bench "step" is one 0.1 ms model step of every trial.  ``value`` is
B * steps * 1e-4 / wall seconds of the timed region (inputs and state
resident in HBM).  The JSON line also carries the B = 1 and B = 64 rates,
the B = 256 : B = 1 ratio,
the roofline of the dominant kernel (per-launch durations from HIP events
on the launch stream) and a CPU baseline: the oracle's torch restatement of
the reference step (csrmv via torch.sparse.mm) on this host's cores, over a
bounded sample of the B = 1 workload.

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
BETA = 26  # algorithmic bytes per neuron-trial-step of the dense update


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed(fn, sync):
    sync()
    t0 = time.perf_counter()
    fn()
    sync()
    return time.perf_counter() - t0


def kernel_times(runner, windows):
    """Average durations (ms) of the drive launch and of the road's
    launches -- push: update, scatter (one fused launch below 64 trials);
    pull: the fused window kernel --
    over ``windows`` eager windows, bracketed by HIP events on the launch
    stream."""
    import torch

    s = runner.sim
    W = runner.window
    stages = [] if runner.fused else [
        ("drive", lambda: runner.table is not None
         and s.drive_table(runner.table, W))]
    if s.road == "pull":
        stages.append(("pull", lambda: s.pull(W)))
    elif s.bstride < 64:
        stages.append(("update_scatter", lambda: s.update_scatter(W)))
    else:
        stages += [("update", lambda: s.update(W)),
                   ("scatter", lambda: s.scatter())]
    ev = {k: [] for k, _ in stages}
    stream = torch.cuda.current_stream()
    # untimed windows first: after the host-side pause that follows the
    # timed region the first eager windows run slow (956 -> 853 us over 20
    # windows of k_pull at B = 256, in a kernel trace) while the GPU ramps
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
    return {k: sum(a.elapsed_time(b) for a, b in v) / len(v)
            for k, v in ev.items()}


def scatter_bytes(conn, raster_rows):
    """Algorithmic bytes of the scatter for given spikes: per spiking
    (neuron, step) row 8 B offsets + 6 B per edge, plus an 8 B int32
    read-modify-write per edge and spiking trial."""
    import numpy as np

    deg = conn.degree()
    rows = 0
    edges_trials = 0
    for sp in raster_rows:  # bool [batch, n] per step
        any_t = sp.any(axis=0)
        rows += int((8 + 6 * deg[any_t]).sum())
        edges_trials += int((deg[None, :] * sp).sum())
    return rows + 8 * edges_trials


def serve_bench(net, conn, args):
    """Serving workload: Poisson arrivals of simulation requests
    (500-20,000 steps, 20-200 driven neurons at 50-250 Hz, 0-50 silenced,
    own seed) served by continuous batching in resident slots, and the same
    requests run back to back by the plain B = 1 engine.  Both in real
    time; results must be identical request by request."""
    import numpy as np

    from flybrain_amd.serve import Server, SimRequest, poisson_requests

    reqs = poisson_requests(conn.n, args.serve_requests, args.serve_rate,
                            seed=1)

    def copies():
        return [SimRequest(**vars(r)) for r in reqs]

    warm = [SimRequest(rid=-1 - k, steps=200, drive=reqs[k].drive,
                       rate_hz=reqs[k].rate_hz, seed=k) for k in range(2)]
    out = {"workload": "poisson_requests", "rate_req_s": args.serve_rate,
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
        log("serve %s: %s" % (name, st))
        del srv
    a, b = runs["batched"], runs["b1_back_to_back"]
    out["results_identical"] = bool(
        a.keys() == b.keys() and all(np.array_equal(a[k], b[k]) for k in a))
    return out


def profiled_traffic(kernel):
    """HBM bytes per launch of ``kernel`` at its largest grid, from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE summaries committed in profiles/
    (KiB per dispatch, scaled by what the same counters read for a plain
    256 MiB copy: FETCH x 2.0, WRITE x 1.0).  None when not profiled."""
    import glob

    best = {}
    for path in sorted(glob.glob(str(ROOT / "profiles" / "*_hbm_counters.md"))):
        for line in open(path):
            parts = line.split()
            if len(parts) < 5 or parts[-3] not in ("FETCH_SIZE", "WRITE_SIZE"):
                continue
            head = line[:60]
            if (kernel + "(" not in head and kernel + "<" not in head
                    and not (kernel.endswith(">") and kernel in head)):
                continue
            grid, name, val = int(parts[-4]), parts[-3], float(parts[-1])
            if grid >= best.get(name, (0, 0))[0]:
                best[name] = (grid, val)
    if "FETCH_SIZE" not in best or "WRITE_SIZE" not in best:
        return None
    return int(1024 * (2.0 * best["FETCH_SIZE"][1]
                       + 1.0 * best["WRITE_SIZE"][1]))


def cpu_baseline(conn, wl, seconds):
    """The oracle's torch restatement of the reference step on this host."""
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
    while time.perf_counter() - t0 < seconds:
        bits[:] = 0
        kick = wl.drive_neuron[rng.random(wl.drive_neuron.size) < p]
        bits[0, kick] = 1
        eng.step(bits)
        steps += 1
    wall = time.perf_counter() - t0
    return {"value": round(steps * 1e-4 / wall, 5), "unit": "trial-s/s",
            "cores": cores, "kind": "port",
            "sample": "%d steps of the B=1 workload (138,639 neurons, 20 "
                      "neurons at 150 Hz), torch CPU csrmv restatement of "
                      "the reference step, %.1f s" % (steps, wall),
            "ms_per_step": round(1e3 * wall / steps, 3)}


def summarize(paths):
    """Per-kernel summary of rocprofv3 databases (kernel trace or --pmc).

    Dispatches are grouped by kernel name and grid size (which tells the
    B = 256 and B = 1 workloads apart).  For counter databases the mean
    value per dispatch is printed next to what the same counter read for
    the 256 MiB calibration copy (``--calib-copy``).
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
            # the median too: a run's quieter first windows (state from
            # rest) pull the mean below the steady-state duration
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


def _short(name):
    name = name.replace("(anonymous namespace)::", "")
    return name[:60]


def main():
    if len(sys.argv) > 2 and sys.argv[1] == "--summarize":
        summarize(sys.argv[2:])
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--b1-steps", type=int, default=10_000)
    ap.add_argument("--b64-steps", type=int, default=10_000)
    ap.add_argument("--graph-windows", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=("all", "b256", "b1", "b64"),
                    default="all")
    ap.add_argument("--serve-requests", type=int, default=128,
                    help="requests of the serving run (0: skip it)")
    ap.add_argument("--serve-rate", type=float, default=200.0,
                    help="Poisson arrival rate of requests per second")
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
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a 256 MiB elementwise copy (counter "
                         "calibration: known bytes read and written)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the whole batch on "
              "one MI355X (trials shard across GPUs with no exchange; see "
              "DESIGN.md).  --gpus %d is not run." % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from flybrain_amd import connectome as C
    from flybrain_amd import workloads as WL
    from flybrain_amd.batch import BatchRunner
    from flybrain_amd.device import DeviceNet

    sync = torch.cuda.synchronize
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

    res = {}
    for name in (("b256", "b1", "b64") if args.only == "all"
                 else (args.only,)):
        if name == "b256":
            wl = WL.sensory_b256(conn.n, seed=0, steps=args.warmup + args.steps)
            steps = args.steps
        elif name == "b64":  # per-trial drive, autotuned trial-block width
            wl = WL.trials_b64(conn.n, seed=0,
                               steps=args.warmup + args.b64_steps)
            steps = args.b64_steps
        else:
            wl = WL.plain_b1(conn.n, seed=0, steps=args.warmup + args.b1_steps)
            steps = args.b1_steps
        tuned = None
        if args.road == "tuned":
            from flybrain_amd.autotune import tune
            t1 = time.perf_counter()
            tuned = tune(net, wl)
            log("%s: tuned %s in %.1f s, ms per window %s"
                % (name, tuned["best"], time.perf_counter() - t1,
                   tuned["times"]))
        r = BatchRunner(net, wl, counts=True, road=args.road)
        if not args.no_graph:
            r.capture(args.graph_windows)
        r.advance(args.warmup, use_graph=not args.no_graph)
        wall = timed(lambda: r.advance(steps, use_graph=not args.no_graph),
                     sync)
        spikes = int(r.counts().sum())
        kt = kernel_times(r, 20)
        res[name] = dict(wl=wl, runner=r, wall=wall, steps=steps,
                         spikes=spikes, kt=kt, tuned=tuned,
                         tps=wl.batch * steps * 1e-4 / wall)
        log("%s: B=%d %d steps %.4f s  %.3f us/step  %.2f trial-s/s  "
            "spikes/trial-step %.2f  kernels(ms) %s"
            % (name, wl.batch, steps, wall, 1e6 * wall / steps,
               res[name]["tps"],
               spikes / (wl.batch * (steps + args.warmup)),
               {k: round(v, 5) for k, v in kt.items()}))

    if args.calib_copy:
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        b = a + 1.0  # reads 256 MiB, writes 256 MiB
        sync()
        log("calibration copy: 256 MiB read + 256 MiB written (%s)"
            % float(b[0]))
        del a, b
    head = res.get("b256") or res.get("b1") or res.get("b64")
    wl = head["wl"]
    n, B, W = conn.n, wl.batch, head["runner"].window
    kt = head["kt"]
    dom = max((k for k in kt if k != "drive"), key=lambda k: kt[k])
    alg = 0
    if dom in ("update", "pull", "update_scatter"):
        alg += BETA * n * B * W
    if dom in ("scatter", "pull", "update_scatter"):
        # spikes of the last window, from a short raster run
        from flybrain_amd.batch import BatchRunner as _BR
        rr = _BR(net, wl, counts=False, raster_len=W)
        rr.advance(min(wl.steps, args.warmup + head["steps"]) // W * W - W)
        rr.advance(W, use_graph=False)
        rows = rr.sim.read_raster(rr.t - W, rr.t)
        alg += scatter_bytes(conn, rows)
        del rr
    achieved = alg / (kt[dom] * 1e-3) / 1e9
    kname = {"update": "k_update", "update_scatter": "k_update",
             "pull": "k_pull"}.get(dom) or (
        "k_scatter_steps" if head["runner"].sim.bstride >= 64 else "k_scatter")
    if dom == "pull":  # the instantiation that ran: window, stride, tile
        hs = head["runner"].sim
        kname = "k_pull<%d, %d, %d>" % (20, hs.bstride, hs.pull_tile)
    traffic = profiled_traffic(kname)
    out.update({
        "value": round(head["tps"], 3),
        "steps": head["steps"], "warmup": args.warmup,
        "ms_per_step": round(1e3 * head["wall"] / head["steps"], 6),
        "config": {"workload": wl.name, "neurons": n,
                   "edges": conn.n_edges, "batch": B,
                   "road": head["runner"].sim.road,
                   "pull_tile": (head["runner"].sim.pull_tile
                                 if head["runner"].sim.road == "pull"
                                 else None),
                   "autotuned": (head["tuned"]["times"] if head["tuned"]
                                 else None),
                   "window_steps": W, "graph_windows": args.graph_windows,
                   "connectome": conn.fingerprint()},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_gbs": (round(traffic / (kt[dom] * 1e-3) / 1e9, 1)
                                     if traffic else None),
                     "algorithmic_bytes_per_launch": int(alg),
                     "avg_launch_ms": round(kt[dom], 5)},
        "kernels_ms": {k: round(v, 5) for k, v in kt.items()},
    })
    if "b1" in res and "b256" in res:
        out["b1"] = {"value": round(res["b1"]["tps"], 3),
                     "steps": res["b1"]["steps"],
                     "ms_per_step": round(1e3 * res["b1"]["wall"]
                                          / res["b1"]["steps"], 6),
                     "kernels_ms": {k: round(v, 5)
                                    for k, v in res["b1"]["kt"].items()}}
        out["ratio_b256_over_b1"] = round(res["b256"]["tps"]
                                          / res["b1"]["tps"], 3)
    if "b64" in res:  # 64 trials, each with its own 20-200 driven neurons
        r64 = res["b64"]
        out["b64"] = {"workload": r64["wl"].name,
                      "value": round(r64["tps"], 3), "steps": r64["steps"],
                      "ms_per_step": round(1e3 * r64["wall"]
                                           / r64["steps"], 6),
                      "road": r64["runner"].sim.road,
                      "pull_tile": (r64["runner"].sim.pull_tile
                                    if r64["runner"].sim.road == "pull"
                                    else None),
                      "autotuned": (r64["tuned"]["times"] if r64["tuned"]
                                    else None),
                      "kernels_ms": {k: round(v, 5)
                                     for k, v in r64["kt"].items()}}
    for name, rr in res.items():
        out.setdefault("spikes_per_trial_step", {})[name] = round(
            rr["spikes"] / (rr["wl"].batch * (rr["steps"] + args.warmup)),
            3)
    if args.serve_requests > 0 and args.only == "all":
        sv = serve_bench(net, conn, args)
        out["serving"] = sv
        out["p50_ms"] = sv["batched"]["p50_ms"]
        out["p99_ms"] = sv["batched"]["p99_ms"]
    if not args.no_cpu:
        b1 = WL.plain_b1(conn.n, seed=0)
        out["cpu_baseline"] = cpu_baseline(conn, b1, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
