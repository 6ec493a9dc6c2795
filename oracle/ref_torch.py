"""The oracle's torch-CPU restatement of the reference step.

TEST INFRASTRUCTURE ONLY: used by the tests (pinned against the golden
vectors) and as the ``cpu_baseline`` leg of bench.py, never by the engine.

It restates Engine._core (lab/flybrain/sim.py:863-953) op for op with
torch CPU tensors, the SpMV as ``torch.sparse.mm`` over the CSR of
W[post, pre] (the reference's csrmv road, sim.py:910-912), so its speed is
the speed of the reference's own CPU path.  The drive is given per step as
the Bernoulli outcomes (a tape), as Engine._poisson would produce them
(sim.py:858-861).
"""

from __future__ import annotations

import numpy as np
import torch

DT = 0.1
PARAMS = {"tauSyn": 5.0, "tDelay": 1.8, "v0": -52.0, "vReset": -52.0,
          "vRest": -52.0, "vThreshold": -45.0, "tauMem": 20.0,
          "tRefrac": 2.2, "scalePoisson": 250, "wScale": 0.275}


class TorchRef:
    def __init__(self, conn, batch=1, no_refractory=None, params=None):
        p = dict(PARAMS)
        if params:
            p.update(params)
        n = conn.n
        pre = np.repeat(np.arange(n, dtype=np.int64), conn.degree())
        idx = torch.from_numpy(np.vstack([conn.cols.astype(np.int64), pre]))
        w = torch.sparse_coo_tensor(idx, torch.from_numpy(
            conn.vals.astype(np.float32)), (n, n)).coalesce()
        self.w_csr = w.to_sparse_csr()
        self.n, self.batch = n, batch
        self.tau_syn_f = DT / p["tauSyn"]
        self.tau_mem_f = DT / p["tauMem"]
        self.poisson_gain = p["scalePoisson"]
        self.w_scale = p["wScale"]
        self.v_rest, self.v_reset = p["vRest"], p["vReset"]
        self.v_th = p["vThreshold"]
        self.refrac_steps = torch.full((n,), float(round(p["tRefrac"] / DT)))
        if no_refractory is not None and len(no_refractory):
            self.refrac_steps[torch.as_tensor(np.asarray(no_refractory,
                                                         np.int64))] = 0.0
        self.buf_len = int(p["tDelay"] / DT) + 1
        self.conductance = torch.zeros(batch, n)
        self.delay_buffer = torch.zeros(batch, self.buf_len, n)
        self.ptr = 0
        self.v = torch.full((batch, n), p["v0"])
        self.spikes = torch.zeros(batch, n)
        self.refrac = self.refrac_steps.unsqueeze(0).repeat(batch, 1).clone()
        self.m_fire = torch.ones(batch, n)
        self.m_out = torch.ones(batch, n)
        self.m_in = torch.ones(batch, n)
        self.m_hold = torch.ones(batch, n)
        self.v_hold = torch.zeros(batch, n)

    def poisson(self, bits):
        bern = torch.as_tensor(bits, dtype=torch.float32)
        return self.w_scale * (bern * self.poisson_gain)

    def step(self, bits):
        voltage_stim = self.poisson(bits)
        recurrent = self.w_scale * torch.sparse.mm(
            self.w_csr, (self.spikes * self.m_out).t().contiguous()).t()
        self.refrac.add_(1.0).mul_(1.0 - self.spikes)
        gate = (self.refrac >= self.refrac_steps.unsqueeze(0)).float()
        delayed = self.delay_buffer[:, self.ptr, :] * self.m_in
        conductance_new = (self.conductance * (1.0 - self.tau_syn_f)
                           + delayed * gate)
        self.delay_buffer[:, self.ptr, :] = recurrent
        self.ptr = (self.ptr + 1) % self.buf_len
        self.v.add_(voltage_stim)
        self.v.add_(self.tau_mem_f * (self.conductance
                                      - (self.v - self.v_rest)))
        self.v.mul_(self.m_hold).add_(self.v_hold)
        self.spikes.copy_(self.v > self.v_th)
        self.spikes.mul_(self.m_fire)
        self.v.sub_((self.v - self.v_reset) * self.spikes)
        self.conductance.copy_(conductance_new - conductance_new * self.spikes)
        return self.spikes
