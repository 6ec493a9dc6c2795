"""Connectome storage for the LIF engine: synthetic generator and the
reference's weight-matrix format.

The engine keeps the synapse matrix by PRESYNAPTIC neuron (the transpose of
the reference's ``W[post, pre]``): row ``pre`` lists every outgoing edge of
that neuron, so one spike reads exactly one contiguous row.  Layout:

* ``offsets`` int64 [n+1]  -- row starts (``offsets[n] == n_edges``);
* ``cols``    int32 [E]    -- postsynaptic neuron of each edge, sorted per row;
* ``vals``    int16 [E]    -- signed synapse count (``sign(pre) * count``).

The weight in millivolts is ``vals * wScale`` and the engine applies
``wScale`` once, after the exact integer sum, as the reference does
(``lab/flybrain/sim.py:906-918``).

Reference format (``lab/flybrain/data.py:221-259``): a torch sparse COO
``W[post, pre]`` with float32 values ``Excitatory x Connectivity``, cached as
``{"indices": int64 [2, E] (post, pre), "values": float32 [E], "shape": (n, n)}``
and coalesced (duplicate entries summed).  :func:`to_reference_coo` and
:func:`from_reference_coo` convert both ways.
"""

from __future__ import annotations

import hashlib
from dataclasses import dataclass, field

import numpy as np

# Largest integer below which any float32 sum of integers is exact in any
# order (lab/flybrain/evspmv.py:108).
F32_EXACT_INT = 1 << 24


@dataclass
class Connectome:
    """Synapse matrix by presynaptic neuron plus optional group labels."""

    n: int
    offsets: np.ndarray  # int64 [n+1]
    cols: np.ndarray  # int32 [E]
    vals: np.ndarray  # int16 [E]
    groups: np.ndarray | None = None  # int16 [n]
    meta: dict = field(default_factory=dict)

    def __post_init__(self):
        self.offsets = np.ascontiguousarray(self.offsets, dtype=np.int64)
        self.cols = np.ascontiguousarray(self.cols, dtype=np.int32)
        self.vals = np.ascontiguousarray(self.vals, dtype=np.int16)
        self.check()

    @property
    def n_edges(self) -> int:
        return int(self.offsets[-1])

    def degree(self) -> np.ndarray:
        return np.diff(self.offsets)

    def check(self):
        """Bounds of every index the kernels dereference, checked once.

        Mirrors ``Events._check_bounds`` (lab/flybrain/evspmv.py:1057-1113):
        an index past the end would not fault reliably inside a caching
        allocator's arena, it would silently read or write a neighbour.
        """
        bad = []
        n = self.n
        if self.offsets.shape != (n + 1,):
            bad.append("offsets has %d entries, expected n+1 = %d"
                       % (self.offsets.size, n + 1))
        else:
            if self.offsets[0] != 0:
                bad.append("offsets[0] = %d, expected 0" % self.offsets[0])
            if np.any(np.diff(self.offsets) < 0):
                bad.append("offsets decrease somewhere")
            if int(self.offsets[-1]) != self.cols.size:
                bad.append("offsets[n] = %d but %d edges"
                           % (int(self.offsets[-1]), self.cols.size))
        if self.cols.size != self.vals.size:
            bad.append("%d cols but %d vals" % (self.cols.size, self.vals.size))
        if self.cols.size:
            lo, hi = int(self.cols.min()), int(self.cols.max())
            if lo < 0 or hi >= n:
                bad.append("post index range [%d, %d] outside [0, %d)"
                           % (lo, hi, n))
        if self.cols.size >= (1 << 31):
            bad.append("more than 2^31 edges")
        if bad:
            raise ValueError("connectome out of bounds:\n  " + "\n  ".join(bad))

    def fingerprint(self) -> str:
        """First 16 hex digits of sha256 over the three CSR arrays."""
        h = hashlib.sha256()
        for a in (self.offsets, self.cols, self.vals):
            h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()[:16]

    def exactness(self) -> dict:
        """Preconditions for an order-free exact integer sum.

        Same three numbers as ``exactness`` (lab/flybrain/evspmv.py:957-994):
        values are integers (always true for int16 storage), spikes are 0/1,
        and the largest sum of |w| over the inputs of one postsynaptic neuron
        stays below 2^24, so the reference's float32 csrmv is exact too.
        """
        absw = np.abs(self.vals.astype(np.int64))
        per_post = np.bincount(self.cols, weights=absw, minlength=self.n)
        worst = float(per_post.max()) if self.n else 0.0
        return {
            "integral": True,
            "min_value": int(self.vals.min()) if self.vals.size else 0,
            "max_value": int(self.vals.max()) if self.vals.size else 0,
            "max_row_abs_sum": worst,
            "f32_exact_limit": F32_EXACT_INT,
            "margin": round(F32_EXACT_INT / worst, 1) if worst else None,
            "ok": bool(worst < F32_EXACT_INT),
        }

    def fan(self):
        """(n_in, n_out) edge counts per neuron (lab/flybrain/sim.py:573-592)."""
        n_out = self.degree().astype(np.int64)
        n_in = np.bincount(self.cols, minlength=self.n).astype(np.int64)
        return n_in, n_out

    def dense_counts(self) -> np.ndarray:
        """Dense int64 ``W[post, pre]`` -- for tiny graphs in tests only."""
        w = np.zeros((self.n, self.n), dtype=np.int64)
        pre = np.repeat(np.arange(self.n), self.degree())
        np.add.at(w, (self.cols.astype(np.int64), pre), self.vals.astype(np.int64))
        return w

    def silence(self, silenced) -> "Connectome":
        """Drop every edge whose pre OR post neuron is silenced.

        Same rule as ``_silence`` (lab/flybrain/sim.py:324-335).
        """
        sil = np.asarray(silenced if silenced is not None else [], dtype=np.int64)
        if sil.size == 0:
            return self
        if sil.min() < 0 or sil.max() >= self.n:
            raise ValueError("silenced index outside [0, %d)" % self.n)
        mask = np.zeros(self.n, dtype=bool)
        mask[sil] = True
        pre = np.repeat(np.arange(self.n), self.degree())
        keep = ~(mask[pre] | mask[self.cols])
        deg = np.bincount(pre[keep], minlength=self.n)
        off = np.zeros(self.n + 1, dtype=np.int64)
        np.cumsum(deg, out=off[1:])
        return Connectome(self.n, off, self.cols[keep], self.vals[keep],
                          self.groups, dict(self.meta, silenced=int(sil.size)))


# ---------------------------------------------------------------- formats


def from_edges(n, pre, post, val, groups=None, meta=None) -> Connectome:
    """Build the by-pre CSR from an edge list; duplicates are summed.

    Summing duplicates is what ``coalesce()`` does to the reference's COO
    (lab/flybrain/data.py:254-258).  Edges whose summed value is zero are
    kept, as coalesce keeps them.
    """
    pre = np.asarray(pre, dtype=np.int64)
    post = np.asarray(post, dtype=np.int64)
    val = np.asarray(val)
    if pre.size and (pre.min() < 0 or pre.max() >= n or post.min() < 0
                     or post.max() >= n):
        raise ValueError("edge index outside [0, %d)" % n)
    vf = val.astype(np.float64)
    if vf.size and np.any(vf != np.round(vf)):
        raise ValueError("synapse values must be integers (signed counts)")
    key = pre * n + post
    order = np.argsort(key, kind="stable")
    key = key[order]
    vi = np.round(vf[order]).astype(np.int64)
    uniq, start = np.unique(key, return_index=True)
    summed = np.add.reduceat(vi, start) if vi.size else vi
    if summed.size and np.abs(summed).max() > 32767:
        raise ValueError("a summed synapse value exceeds the int16 range")
    upre = uniq // n
    upost = uniq % n
    off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(upre, minlength=n), out=off[1:])
    return Connectome(n, off, upost.astype(np.int32), summed.astype(np.int16),
                      groups, dict(meta or {}))


def to_reference_coo(conn: Connectome) -> dict:
    """The reference's cached weight format (lab/flybrain/data.py:230-236).

    ``indices`` rows are (post, pre), ``values`` float32 signed counts,
    ``shape`` (n, n).  Entries come out sorted by (post, pre), which is
    the order ``coalesce()`` produces.
    """
    import torch

    pre = np.repeat(np.arange(conn.n, dtype=np.int64), conn.degree())
    post = conn.cols.astype(np.int64)
    order = np.lexsort((pre, post))
    idx = np.vstack([post[order], pre[order]])
    return {
        "indices": torch.from_numpy(np.ascontiguousarray(idx)),
        "values": torch.from_numpy(conn.vals[order].astype(np.float32)),
        "shape": (conn.n, conn.n),
    }


def from_reference_coo(obj, n=None) -> Connectome:
    """Read the reference format: a dict as above or a torch sparse COO."""
    import torch

    if isinstance(obj, dict):
        idx = torch.as_tensor(obj["indices"]).cpu().numpy()
        val = torch.as_tensor(obj["values"]).cpu().numpy()
        shape = tuple(int(s) for s in obj["shape"])
    else:
        w = obj.coalesce().cpu()
        idx = w.indices().numpy()
        val = w.values().numpy()
        shape = tuple(int(s) for s in w.shape)
    if shape[0] != shape[1]:
        raise ValueError("weight matrix is not square: %s" % (shape,))
    n = shape[0] if n is None else int(n)
    return from_edges(n, idx[1], idx[0], val)


# FlyWire v783 tables as the reference reads them (lab/flybrain/data.py:
# 41-42, 239-252): the connectivity parquet's index and value columns, and
# the completeness CSV, whose row count is N.
FLYWIRE_COLUMNS = ("Presynaptic_Index", "Postsynaptic_Index",
                   "Excitatory x Connectivity")


def from_flywire_tables(conn_parquet, completeness_csv=None, n=None
                        ) -> Connectome:
    """Build the connectome from the reference's FlyWire files: edges from
    the connectivity parquet's three columns above (duplicates summed, as
    the reference's ``coalesce()``), N from the completeness CSV's rows
    (``pd.read_csv(..., index_col=0)``, data.py:239-240) or ``n``."""
    import pandas as pd

    if n is None:
        if completeness_csv is None:
            raise ValueError("N comes from the completeness CSV or n=")
        n = len(pd.read_csv(completeness_csv, index_col=0))
    df = pd.read_parquet(conn_parquet, columns=list(FLYWIRE_COLUMNS))
    return from_edges(int(n), df[FLYWIRE_COLUMNS[0]].to_numpy(np.int64),
                      df[FLYWIRE_COLUMNS[1]].to_numpy(np.int64),
                      df[FLYWIRE_COLUMNS[2]].to_numpy(np.float32))


def to_flywire_tables(conn: Connectome, conn_parquet, completeness_csv=None):
    """Write ``conn`` in the reference's FlyWire table layout (the three
    connectivity columns; a completeness CSV of N rows indexed by neuron),
    so a synthetic connectome goes through the reference's loader."""
    import pandas as pd

    pre = np.repeat(np.arange(conn.n, dtype=np.int64), conn.degree())
    pd.DataFrame({FLYWIRE_COLUMNS[0]: pre,
                  FLYWIRE_COLUMNS[1]: conn.cols.astype(np.int64),
                  FLYWIRE_COLUMNS[2]: conn.vals.astype(np.int64)}).to_parquet(
        conn_parquet, index=False)
    if completeness_csv is not None:
        pd.DataFrame({"Completed": np.ones(conn.n, bool)},
                     index=pd.Index(np.arange(conn.n), name="root_id")
                     ).to_csv(completeness_csv)


def default_paths(root=None):
    """Where the reference keeps its connectome (lab/flybrain/data.py:
    31-48): ``cache/weights_coo.pt`` and ``upstream/data/2025_*_783``
    under its project root -- ``root``, else ``$FLYBRAIN_ROOT``, else the
    current directory."""
    import os
    from pathlib import Path

    base = Path(root or os.environ.get("FLYBRAIN_ROOT") or ".")
    return {"cache": base / "cache" / "weights_coo.pt",
            "connectivity": base / "upstream" / "data"
            / "2025_Connectivity_783.parquet",
            "completeness": base / "upstream" / "data"
            / "2025_Completeness_783.csv"}


def load_default(root=None) -> Connectome:
    """The reference's default connectome (data.load_weights,
    lab/flybrain/data.py:221-259): its COO cache if present (loaded with
    ``weights_only=True``), else the FlyWire connectivity table and the
    completeness CSV.  Raises FileNotFoundError naming the paths when
    neither is there."""
    p = default_paths(root)
    if p["cache"].exists():
        import torch

        return from_reference_coo(torch.load(p["cache"], weights_only=True))
    if p["connectivity"].exists() and p["completeness"].exists():
        return from_flywire_tables(p["connectivity"], p["completeness"])
    raise FileNotFoundError(
        "no connectome given and none at the reference's default locations "
        "(%s, or %s with %s); pass weights, or set FLYBRAIN_ROOT"
        % (p["cache"], p["connectivity"], p["completeness"]))


def save(conn: Connectome, path):
    """Write the by-pre CSR as an uncompressed .npz."""
    extra = {} if conn.groups is None else {"groups": conn.groups}
    np.savez(path, n=np.int64(conn.n), offsets=conn.offsets, cols=conn.cols,
             vals=conn.vals, **extra)


def load(path) -> Connectome:
    z = np.load(path, allow_pickle=False)
    groups = z["groups"] if "groups" in z.files else None
    return Connectome(int(z["n"]), z["offsets"], z["cols"], z["vals"], groups)


# -------------------------------------------------------------- generator


def _powerlaw_table(exponent, kmax):
    k = np.arange(1, kmax + 1, dtype=np.float64)
    p = k ** (-float(exponent))
    cdf = np.cumsum(p)
    return cdf / cdf[-1]


def _degrees(rng, n, n_edges, mean, sigma, max_deg):
    """Log-normal out-degrees rescaled so they sum to ``n_edges`` exactly."""
    mu = np.log(mean) - 0.5 * sigma * sigma
    d = np.rint(rng.lognormal(mu, sigma, size=n))
    d = np.clip(d, 1, max_deg)
    if n_edges is None:
        return d.astype(np.int64)
    if not (n <= n_edges <= n * max_deg):
        raise ValueError("cannot place %d edges on %d neurons" % (n_edges, n))
    d = np.clip(np.floor(d * (n_edges / d.sum())), 1, max_deg).astype(np.int64)
    while True:
        diff = int(n_edges - d.sum())
        if diff == 0:
            return d
        if diff > 0:
            room = np.nonzero(d < max_deg)[0]
            pick = rng.choice(room, size=min(diff, room.size), replace=False)
            d[pick] += 1
        else:
            room = np.nonzero(d > 1)[0]
            pick = rng.choice(room, size=min(-diff, room.size), replace=False)
            d[pick] -= 1


def _sample_unique(rng, rows, k, draw):
    """Per row r, ``k[r]`` distinct values from ``draw(row_ids) -> values``.

    Draws with replacement, then redraws duplicates until every row's
    values are distinct.  ``draw`` must sample uniformly from a pool larger
    than ``k`` for the row (callers route near-full rows elsewhere).
    """
    owner = np.repeat(rows, k)
    vals = draw(owner)
    while True:
        key = owner * (1 << 32) + vals
        order = np.argsort(key, kind="stable")
        sk = key[order]
        dup = np.zeros(sk.size, dtype=bool)
        dup[1:] = sk[1:] == sk[:-1]
        if not dup.any():
            return owner, vals
        redo = order[dup]
        vals[redo] = draw(owner[redo])


def synthetic(n, n_edges=None, *, mean_out=109.0, sigma=1.0, count_exp=2.0,
              count_max=2000, p_excitatory=0.55, n_groups=36, p_within=0.8,
              group_sigma=1.0, seed=0) -> Connectome:
    """Seeded synthetic connectome with the statistics of the benchmark.

    * out-degree ~ round(LogNormal(ln(mean_out) - sigma^2/2, sigma)), clipped
      to [1, n-1] and rescaled so the total is exactly ``n_edges``;
    * targets are distinct, never the neuron itself;
    * with ``n_groups > 1`` neurons fall into groups of log-normal size and a
      fraction ``p_within`` of all edges stays inside the presynaptic
      neuron's group (rows of groups too small for their share spill out,
      the other rows keep slightly more inside to compensate);
      group membership is a seeded random permutation of neuron ids, so a
      group is not a contiguous id range;
    * synapse count ~ discrete power law P(k) ~ k^-count_exp on [1, count_max];
    * sign per presynaptic neuron, excitatory with probability p_excitatory.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    n = int(n)
    max_deg = n - 1
    deg = _degrees(rng, n, n_edges, mean_out, sigma, max_deg)

    if n_groups and n_groups > 1:
        gs = rng.lognormal(0.0, group_sigma, size=n_groups)
        size = np.maximum(1, np.floor(gs / gs.sum() * n)).astype(np.int64)
        size[np.argmax(size)] += n - size.sum()
    else:
        size = np.array([n], dtype=np.int64)
        p_within = 1.0
    start = np.zeros(size.size + 1, dtype=np.int64)
    np.cumsum(size, out=start[1:])
    # position p in [0, n) is group-sorted; perm maps position -> neuron id
    gpos = np.repeat(np.arange(size.size), size)
    perm = rng.permutation(n)

    gsz = size[gpos]

    def within(p):
        return np.minimum(np.rint(deg * p).astype(np.int64), gsz - 1)

    # rows of small groups cannot keep p_within of their edges inside; raise
    # the share of the others so the graph as a whole has p_within inside
    target = int(round(p_within * deg.sum()))
    lo, hi = float(p_within), 1.0
    if within(hi).sum() <= target:
        lo = hi
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        if within(mid).sum() < target:
            lo = mid
        else:
            hi = mid
    k_in = within(lo)
    k_out = deg - k_in
    cap_out = n - gsz
    over = k_out > cap_out
    k_in[over] += k_out[over] - cap_out[over]
    k_out = deg - k_in

    rows = np.arange(n, dtype=np.int64)
    g0 = start[gpos]

    def draw_in(owner):
        # uniform over the owner's group, skipping the owner itself
        u = (rng.random(owner.size) * (gsz[owner] - 1)).astype(np.int64)
        local = owner - g0[owner]
        u = u + (u >= local)
        return g0[owner] + u

    def draw_out(owner):
        # uniform over positions outside the owner's group
        u = (rng.random(owner.size) * (n - gsz[owner])).astype(np.int64)
        return u + np.where(u >= g0[owner], gsz[owner], 0)

    parts_pre, parts_post = [], []
    for k, pool, draw in ((k_in, gsz - 1, draw_in), (k_out, n - gsz, draw_out)):
        dense = (k > 0) & (k * 3 > pool)  # near-full rows: exact draw per row
        easy = (k > 0) & ~dense
        o, v = _sample_unique(rng, rows[easy], k[easy], draw)
        parts_pre.append(o)
        parts_post.append(v)
        for r in np.nonzero(dense)[0]:
            if draw is draw_in:
                cand = np.arange(g0[r], g0[r] + gsz[r])
                cand = cand[cand != r]
            else:
                cand = np.concatenate([np.arange(0, g0[r]),
                                       np.arange(g0[r] + gsz[r], n)])
            parts_pre.append(np.full(k[r], r, dtype=np.int64))
            parts_post.append(rng.choice(cand, size=k[r], replace=False))
    pre_pos = np.concatenate(parts_pre)
    post_pos = np.concatenate(parts_post)

    cdf = _powerlaw_table(count_exp, count_max)
    count = np.searchsorted(cdf, rng.random(pre_pos.size), side="right") + 1
    count = np.minimum(count, count_max)
    sign_pos = np.where(rng.random(n) < p_excitatory, 1, -1)
    val = sign_pos[pre_pos] * count

    pre_id = perm[pre_pos]
    post_id = perm[post_pos]
    groups = np.empty(n, dtype=np.int16)
    groups[perm] = gpos.astype(np.int16)
    meta = dict(kind="synthetic", n=n, n_edges=int(pre_id.size),
                mean_out=mean_out, sigma=sigma, count_exp=count_exp,
                count_max=count_max, p_excitatory=p_excitatory,
                n_groups=int(size.size), p_within=float(p_within),
                group_sigma=group_sigma, seed=seed)
    conn = from_edges(n, pre_id, post_id, val, groups, meta)
    if n_edges is not None and conn.n_edges != n_edges:
        raise RuntimeError("generator produced %d edges, wanted %d"
                           % (conn.n_edges, n_edges))
    return conn


# Benchmark connectomes (sizes fixed by the workload definitions).
SMALL = dict(n=10_000, n_edges=1_090_000, n_groups=0, seed=1)
FULL = dict(n=138_639, n_edges=15_091_983, n_groups=36, seed=2)


def small(seed=None) -> Connectome:
    """10,000 neurons, 1.09M edges, uniform targets."""
    kw = dict(SMALL)
    if seed is not None:
        kw["seed"] = seed
    return synthetic(**kw)


def full(seed=None) -> Connectome:
    """138,639 neurons, 15,091,983 edges, 36 groups, 80% within-group."""
    kw = dict(FULL)
    if seed is not None:
        kw["seed"] = seed
    return synthetic(**kw)
