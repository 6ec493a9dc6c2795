"""Write golden vectors of the reference LIF engine (run once, on the CPU).

This script imports the reference engine (lab/flybrain/sim.py) from
/root/reference on the CPU, feeds it recorded Poisson tapes instead of
torch's Bernoulli generator, and writes what it computed to small .npz
fixtures next to this file.  Only this script touches the reference; the
fixtures are data (inputs and expected outputs) and the tests read them.

What only running the reference pins, and is therefore recorded here:
  * the float32 result of the torch op sequence of Engine._core
    (sim.py:863-953) -- scalar casts, op order, masks, reset, ring;
  * the readout of run() (sim.py:1040-1117): schedule, bins, counts;
  * the silencing of build_engine (sim.py:324-335, :956-987).

Fixtures:
  tiny_masks.npz  256 neurons, B=2, 600 steps, all four masks switched on
                  and off, per-trial drive, per-step state fingerprints
  small10k.npz    10,000 neurons / 1.09M edges, 20 neurons at 150 Hz,
                  B=1, 1,000 steps (the small CPU workload)
  run_sched.npz   run() with two windowed stimuli, a silenced set, B=2,
                  60 ms in 2.5 ms bins

Run from the repo root:  python tests/golden/make_golden.py
"""

from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parents[1]
REF_LAB = Path("/root/reference/lab")
sys.path.insert(0, str(REPO / "fly-in-silico_amd"))
sys.path.insert(0, str(REF_LAB))

import torch  # noqa: E402

from flybrain_amd import connectome as C  # noqa: E402

torch.set_num_threads(4)

# The tiny graph: heavier synapse-count tail than the benchmark graphs so
# that recurrent spikes, inhibition and refractory gating all occur.
TINY = dict(n=256, n_edges=None, mean_out=40, count_exp=1.2, n_groups=4,
            seed=11)


def fp(a) -> str:
    """First 16 hex digits of sha256 over the raw bytes (probe/_gate.py:50-53)."""
    a = np.ascontiguousarray(a)
    return hashlib.sha256(a.tobytes()).hexdigest()[:16]


def canon(a):
    """Map -0.0 to +0.0 (x + 0.0 does exactly that and nothing else)."""
    return (np.asarray(a, dtype=np.float32) + np.float32(0.0)).astype(np.float32)


def state_fps(eng):
    return [fp(eng.v.numpy()), fp(eng.conductance.numpy()),
            fp(eng.spikes.numpy()), fp(eng.refrac.numpy()),
            fp(canon(eng.delay_buffer.numpy()))]


def ref_modules():
    from flybrain import sim as S  # the reference engine

    return S


def weights_tensor(conn):
    obj = C.to_reference_coo(conn)
    return torch.sparse_coo_tensor(obj["indices"], obj["values"],
                                   obj["shape"]).coalesce()


def tape_poisson(eng, bits_of_step):
    """Replace the Bernoulli draw by a tape, keeping _poisson's ops.

    sim.py:858-861: ``w_scale * (bernoulli(rates * prob_scale) * gain)``.
    Only the Bernoulli outcome changes; the two multiplications stay.
    """
    def _poisson(rates):
        bern = torch.from_numpy(bits_of_step(eng.calls - 1).astype(np.float32))
        poisson = bern * eng.poisson_gain
        return eng.w_scale * poisson
    eng._poisson = _poisson


def tiny_masks(S):
    conn = C.synthetic(**TINY)
    n, B, steps = conn.n, 2, 600
    rng = np.random.Generator(np.random.PCG64(5))
    drv = [np.sort(rng.choice(n, 24, replace=False)) for _ in range(B)]
    bits = np.zeros((steps, B, n), np.uint8)
    for b in range(B):
        bits[:, b, drv[b]] = rng.random((steps, 24)) < 0.05
    no_ref = np.unique(np.concatenate(drv))
    # mask schedule: (step, kwargs for set_masks)
    pick = rng.choice(n, 40, replace=False)
    sched = [
        (100, dict(fire=pick[0:6], out=pick[6:12], inc=pick[12:18],
                   hold=pick[18:24], v_clamp=-60.0)),
        (250, dict(out=pick[24:32])),
        (300, dict(fire=[], out=[], inc=[], hold=[])),
        (400, dict(hold=pick[32:40])),
        (520, dict(hold=[])),
    ]
    eng = S.build_engine(weights=weights_tensor(conn), neurons=np.arange(n),
                         device="cpu", batch=B, seed=0, no_refractory=no_ref,
                         graphs=False, spmv=S.SPMV_CSR)
    tape_poisson(eng, lambda t: bits[t])
    rates = torch.zeros(B, n)
    fps, rec, raster = [], np.zeros((steps, B, n), np.float32), \
        np.zeros((steps, B, n), np.uint8)
    v_snap, g_snap, r_snap = [], [], []
    negzero = 0
    si = 0
    for t in range(steps):
        while si < len(sched) and sched[si][0] == t:
            eng.set_masks(**sched[si][1])
            si += 1
        slot = eng.ptr
        sp = eng.step(rates)
        raw = eng.delay_buffer[:, slot, :].numpy()
        negzero += int(np.sum((raw == 0) & np.signbit(raw)))
        rec[t] = canon(raw)
        raster[t] = sp.numpy().astype(np.uint8)
        fps.append(state_fps(eng))
        if (t + 1) % 50 == 0:
            v_snap.append(eng.v.numpy().copy())
            g_snap.append(eng.conductance.numpy().copy())
            r_snap.append(eng.refrac.numpy().copy())
    sched_steps = np.array([s for s, _ in sched], np.int64)
    sched_kw = {}
    for k, (s, kw) in enumerate(sched):
        for name, val in kw.items():
            sched_kw["sched%d_%s" % (k, name)] = np.asarray(val)
    np.savez_compressed(
        HERE / "tiny_masks.npz",
        conn_offsets=conn.offsets, conn_cols=conn.cols, conn_vals=conn.vals,
        conn_fp=np.array(conn.fingerprint()),
        tape_bits=np.packbits(bits, axis=-1), n=np.int64(n), batch=np.int64(B),
        steps=np.int64(steps), no_refractory=no_ref,
        sched_steps=sched_steps, **sched_kw,
        fps=np.array(fps), rec=rec, raster=np.packbits(raster, axis=-1),
        v_snap=np.array(v_snap), g_snap=np.array(g_snap),
        refrac_snap=np.array(r_snap), negzero_in_rec=np.int64(negzero),
        exactness_max_row=np.float64(
            S_exactness(conn)["макс сумма модулей по строке"]))
    print("tiny_masks: spikes", int(raster.sum()), "negzero", negzero)


def S_exactness(conn):
    from flybrain import evspmv

    return evspmv.exactness(weights_tensor(conn).to_sparse_csr())


def small10k(S):
    conn = C.small()
    n, B, steps = conn.n, 1, 1000
    rng = np.random.Generator(np.random.PCG64(7))
    drv = np.sort(rng.choice(n, 20, replace=False))
    p = np.float32(150.0) * np.float32(1e-4)
    ev_t, ev_i = np.nonzero(rng.random((steps, 20), dtype=np.float32) < p)
    ev_i = drv[ev_i]

    def bits_of(t):
        b = np.zeros((B, n), np.uint8)
        b[0, ev_i[ev_t == t]] = 1
        return b

    eng = S.build_engine(weights=weights_tensor(conn), neurons=np.arange(n),
                         device="cpu", batch=B, seed=0, no_refractory=drv,
                         graphs=False, spmv=S.SPMV_CSR)
    tape_poisson(eng, bits_of)
    rates = torch.zeros(B, n)
    fps, sp_t, sp_i = [], [], []
    rc_t, rc_i, rc_v = [], [], []
    for t in range(steps):
        slot = eng.ptr
        sp = eng.step(rates)
        idx = np.nonzero(sp.numpy()[0])[0]
        sp_t.append(np.full(idx.size, t, np.int32))
        sp_i.append(idx.astype(np.int32))
        rec = eng.delay_buffer[0, slot, :].numpy()
        nz = np.nonzero(rec)[0]
        rc_t.append(np.full(nz.size, t, np.int32))
        rc_i.append(nz.astype(np.int32))
        rc_v.append(rec[nz].copy())
        fps.append(state_fps(eng))
    np.savez_compressed(
        HERE / "small10k.npz",
        conn_fp=np.array(conn.fingerprint()), n=np.int64(n),
        n_edges=np.int64(conn.n_edges), steps=np.int64(steps),
        driven=drv, tape_t=ev_t.astype(np.int32), tape_i=ev_i.astype(np.int32),
        fps=np.array(fps), spike_t=np.concatenate(sp_t),
        spike_i=np.concatenate(sp_i), rec_t=np.concatenate(rc_t),
        rec_i=np.concatenate(rc_i), rec_v=np.concatenate(rc_v),
        v_final=eng.v.numpy().copy(), g_final=eng.conductance.numpy().copy(),
        refrac_final=eng.refrac.numpy().copy())
    print("small10k: spikes", int(sum(a.size for a in sp_i)),
          "drive events", ev_t.size)


def run_sched(S):
    conn = C.synthetic(**TINY)
    n, B = conn.n, 2
    rng = np.random.Generator(np.random.PCG64(9))
    a = np.sort(rng.choice(n, 12, replace=False))
    b = np.sort(rng.choice(np.setdiff1d(np.arange(n), a), 10, replace=False))
    sil = np.sort(rng.choice(np.setdiff1d(np.arange(n), np.union1d(a, b)), 8,
                             replace=False))
    stimuli = [S.Stimulus(indices=a, rate_hz=400.0, t_on_ms=5.0,
                          t_off_ms=32.5, name="a"),
               S.Stimulus(indices=b, rate_hz=900.0, t_on_ms=20.0,
                          t_off_ms=None, name="b")]
    cfg = S.RunConfig(t_run_s=0.06, bin_ms=2.5, n_trials=B, seed=0,
                      device="cpu", speed=S.REFERENCE, spmv=S.SPMV_CSR)
    urng = np.random.Generator(np.random.PCG64(13))
    tape, rates_a, rates_b = [], [], []
    orig = S.Engine._poisson

    def _poisson(self, rates):
        p = (rates * self.prob_scale).numpy()
        u = urng.random(p.shape, dtype=np.float32)
        bits = (u < p).astype(np.float32)
        tape.append(np.packbits(bits.astype(np.uint8), axis=-1))
        rates_a.append(float(rates[0, a[0]]))
        rates_b.append(float(rates[0, b[0]]))
        poisson = torch.from_numpy(bits) * self.poisson_gain
        return self.w_scale * poisson

    S.Engine._poisson = _poisson
    try:
        res = S.run(stimuli, silenced=sil, config=cfg, neurons=np.arange(n),
                    weights=weights_tensor(conn), progress=False)
    finally:
        S.Engine._poisson = orig
    np.savez_compressed(
        HERE / "run_sched.npz",
        conn_fp=np.array(conn.fingerprint()), n=np.int64(n), batch=np.int64(B),
        conn_offsets=conn.offsets, conn_cols=conn.cols, conn_vals=conn.vals,
        stim_a=a, stim_b=b, silenced=sil, t_run_s=np.float64(0.06),
        bin_ms=np.float64(2.5), tape_bits=np.array(tape),
        rates_a=np.array(rates_a, np.float32),
        rates_b=np.array(rates_b, np.float32),
        counts=res.counts, bins=res.bins)
    print("run_sched: total spikes", int(res.counts.sum()), "bins",
          res.bins.shape)


def main():
    S = ref_modules()
    tiny_masks(S)
    small10k(S)
    run_sched(S)


if __name__ == "__main__":
    main()
