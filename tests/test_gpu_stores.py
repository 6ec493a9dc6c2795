"""Plain stores from many workgroups into shared 128-byte lines.

A scatter through a random permutation makes the 32 words (or 128 bytes)
of every line land from 32 (128) unrelated threads of a grid of thousands
of workgroups, spread over all XCDs.  Every value must be where it was
sent, every time.  This checks the hypothesis, recorded while an
out-of-bounds store was still in the engine, that such stores could be
lost; the engine's per-workgroup line layouts do not depend on it.
"""

from __future__ import annotations

import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype,n", [(torch.int32, 1 << 24),
                                     (torch.uint8, 1 << 24),
                                     (torch.int64, 1 << 22)])
def test_scattered_plain_stores_all_arrive(dtype, n):
    g = torch.Generator(device="cuda")
    g.manual_seed(1234)
    vals = (torch.arange(n, device="cuda") % 251 + 1).to(dtype)
    for rep in range(8):
        perm = torch.randperm(n, device="cuda", generator=g)
        out = torch.zeros(n, dtype=dtype, device="cuda")
        out[perm] = vals  # thread i stores vals[i] at perm[i]
        back = out[perm]
        bad = int((back != vals).sum())
        assert bad == 0, "rep %d: %d of %d stores lost" % (rep, bad, n)
