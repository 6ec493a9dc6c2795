"""MI355X engine for the whole-brain LIF step of R1ens/fly-in-silico.

Modules
-------
connectome  synthetic connectome generator, by-presynaptic CSR, reference
            COO format (lab/flybrain/data.py:221-259)
params      model constants (lab/flybrain/sim.py:36-48) and float32 forms
device      device state of a batch of trials and the kernel launches
sim         drop-in Engine / build_engine / run / Result (lab/flybrain/sim.py)
batch       batched trials with per-trial drive, masks and seeds
serve       continuous-batching request loop
"""

__version__ = "0.1.0"
