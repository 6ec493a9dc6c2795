"""Model constants (CPU): the product's float32 constants equal the
oracle's, and a parameter dict is taken whole, as Engine.__init__ takes it
(lab/flybrain/sim.py:350)."""

from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from flybrain_amd import params as P


@pytest.mark.parametrize("t_refrac", [2.2, 0.0, 0.3, 0.7, 12.6])
def test_constants_equal_the_oracles(t_refrac):
    p = dict(P.MODEL_PARAMS, tRefrac=t_refrac)
    got, want = P.derived(p), O.derived(p)
    for k in ("w_scale", "a_syn", "a_mem", "v_rest", "v_reset", "v_th", "v0",
              "gain", "prob_scale"):
        assert np.float32(got[k]).tobytes() == np.float32(want[k]).tobytes(), k
    assert got["refrac_steps"] == int(want["refrac_steps"])
    assert got["buf_len"] == want["buf_len"] == 19


def test_defaults_are_the_references():
    c = P.derived()
    assert c["refrac_steps"] == 22  # round(2.2 / 0.1)
    assert c["a_syn"] == np.float32(0.98) and c["a_mem"] == np.float32(0.005)
    assert c["prob_scale"] == np.float32(1e-4)


def test_partial_dict_raises_like_the_reference():
    with pytest.raises(KeyError):
        P.derived({"tRefrac": 0.3})


def test_refractory_beyond_the_counter_is_refused():
    with pytest.raises(ValueError):
        P.derived(dict(P.MODEL_PARAMS, tRefrac=12.8))
