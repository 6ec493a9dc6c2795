"""Benchmark workloads: per-trial drives, masks and seeds as plain arrays.

A ``Workload`` lists, for ``batch`` independent trials,
  * one 64-bit Philox seed per trial (trial id 0: a trial's kicks depend on
    its seed, neuron and step only, never on its slot in the batch);
  * drive entries (neuron, trial, group) with a rate table
    rate_hz[segment, group] switched every ``seg_steps`` steps; every
    driven neuron is exempt from refractoriness in its own trial, as
    ``run`` exempts stimulated neurons (lab/flybrain/sim.py:1031-1035);
  * m_out mask pairs (neuron, trial): the neuron's spikes are not
    transmitted in that trial (Engine.set_masks(out=...), sim.py:510-571).

The workloads on the 138,639-neuron / 15,091,983-edge graph:
  plain_b1       20 seeded neurons at 150 Hz, B = 1, 10,000 steps
  plain_b256     the plain_b1 drive in 256 trials with distinct seeds
  trials_b64     B = 64, per trial 20-200 neurons at U(50, 250) Hz
  sensory_b256   B = 256, four fixed index sets (1,116 + 1,133 olfactory-
                 type, 410 gustatory-type, 10,582 photoreceptor-type) with
                 a U(0, 300) Hz rate per (trial, set) for every 10 ms;
                 per-trial m_out masks on 0-50 random neurons; 5,000 steps
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

SENSORY_SETS = (("olfactory_L", 1116), ("olfactory_R", 1133),
                ("gustatory", 410), ("photoreceptor", 10582))


@dataclass
class Workload:
    name: str
    batch: int
    steps: int
    seeds: np.ndarray  # uint64 [batch]
    drive_neuron: np.ndarray  # int64 [E]
    drive_trial: np.ndarray  # int64 [E]
    drive_group: np.ndarray  # int64 [E]
    rate_hz: np.ndarray  # float32 [n_seg, n_groups]
    seg_steps: int
    out_neuron: np.ndarray  # int64 [M]
    out_trial: np.ndarray  # int64 [M]

    def trial(self, b) -> "Workload":
        """Trial ``b`` alone, as a B = 1 workload with the same inputs."""
        d = self.drive_trial == b
        o = self.out_trial == b
        groups = np.unique(self.drive_group[d])
        remap = np.full(self.rate_hz.shape[1], -1, np.int64)
        remap[groups] = np.arange(groups.size)
        return Workload(
            "%s[trial %d]" % (self.name, b), 1, self.steps,
            self.seeds[b:b + 1].copy(), self.drive_neuron[d],
            np.zeros(int(d.sum()), np.int64), remap[self.drive_group[d]],
            self.rate_hz[:, groups].copy(), self.seg_steps,
            self.out_neuron[o], np.zeros(int(o.sum()), np.int64))

    @property
    def n_drive(self):
        return int(self.drive_neuron.size)


def _seeds(rng, batch):
    return rng.integers(1, 2**63, size=batch, dtype=np.uint64)


def plain_b1(n, seed=0, steps=10_000, n_driven=20, rate_hz=150.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    idx = np.sort(rng.choice(n, n_driven, replace=False))
    return Workload("plain_b1", 1, steps, _seeds(rng, 1), idx,
                    np.zeros(idx.size, np.int64), np.zeros(idx.size, np.int64),
                    np.array([[rate_hz]], np.float32), 1 << 30,
                    np.zeros(0, np.int64), np.zeros(0, np.int64))


def plain_replicated(n, seed=0, batch=256, steps=10_000, n_driven=20,
                     rate_hz=150.0):
    """The plain_b1 drive (the same 20 neurons at 150 Hz) in ``batch``
    trials, each with its own Philox seed: every trial is the plain
    workload, so its rate is directly comparable with B = 1."""
    one = plain_b1(n, seed=seed, steps=steps, n_driven=n_driven,
                   rate_hz=rate_hz)
    rng = np.random.Generator(np.random.PCG64(seed + 7_000_001))
    idx = one.drive_neuron
    return Workload("plain_b%d" % batch, batch, steps, _seeds(rng, batch),
                    np.tile(idx, batch), np.repeat(np.arange(batch), idx.size),
                    np.zeros(idx.size * batch, np.int64), one.rate_hz,
                    one.seg_steps, np.zeros(0, np.int64),
                    np.zeros(0, np.int64))


def trials_b64(n, seed=0, batch=64, steps=10_000):
    rng = np.random.Generator(np.random.PCG64(seed))
    ne, tr = [], []
    for b in range(batch):
        k = int(rng.integers(20, 201))
        ne.append(np.sort(rng.choice(n, k, replace=False)))
        tr.append(np.full(k, b, np.int64))
    ne = np.concatenate(ne)
    tr = np.concatenate(tr)
    rates = rng.uniform(50.0, 250.0, size=ne.size).astype(np.float32)
    return Workload("trials_b64", batch, steps, _seeds(rng, batch), ne, tr,
                    np.arange(ne.size, dtype=np.int64), rates[None, :],
                    1 << 30, np.zeros(0, np.int64), np.zeros(0, np.int64))


def sensory_sets(n, seed=0):
    """Four disjoint fixed index sets of the sensory sizes."""
    rng = np.random.Generator(np.random.PCG64(seed + 1_000_003))
    total = sum(k for _, k in SENSORY_SETS)
    pick = rng.choice(n, total, replace=False)
    out, at = [], 0
    for _, k in SENSORY_SETS:
        out.append(np.sort(pick[at:at + k]))
        at += k
    return out


def sensory_b256(n, seed=0, batch=256, steps=5_000, seg_ms=10.0,
                 max_rate=300.0, max_silenced=50):
    rng = np.random.Generator(np.random.PCG64(seed))
    sets = sensory_sets(n, seed)
    seg_steps = int(round(seg_ms / 0.1))
    n_seg = max(1, -(-steps // seg_steps))
    ns = len(sets)
    rate = rng.uniform(0.0, max_rate, size=(n_seg, batch * ns)).astype(np.float32)
    ne, tr, gr = [], [], []
    for b in range(batch):
        for k, s in enumerate(sets):
            ne.append(s)
            tr.append(np.full(s.size, b, np.int64))
            gr.append(np.full(s.size, b * ns + k, np.int64))
    on, ot = [], []
    for b in range(batch):
        m = int(rng.integers(0, max_silenced + 1))
        on.append(np.sort(rng.choice(n, m, replace=False)))
        ot.append(np.full(m, b, np.int64))
    return Workload("sensory_b256", batch, steps, _seeds(rng, batch),
                    np.concatenate(ne), np.concatenate(tr), np.concatenate(gr),
                    rate, seg_steps, np.concatenate(on), np.concatenate(ot))
