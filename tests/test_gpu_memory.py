"""Memory planning of the dWaveOp store (GPU test): a batch counts its own
store and checkpoints as available, so repeated gradients keep the segment
length of the first (no recompute appears after the first call)."""
import numpy as np
import pytest

from tests.problems import Synthetic

pytestmark = pytest.mark.gpu


def test_repeated_forward_keeps_the_whole_history(gpu, monkeypatch):
    import torch
    from pysit_amd.device import DeviceBatch
    probs = [Synthetic((60, 70), ((6, 6), (5, 5)), (0.01, 0.01), 120, 8, seed=s) for s in (1, 2)]
    p0 = probs[0]
    b = DeviceBatch([p.geometry() for p in probs], p0.model, p0.deltas, p0.dt, p0.order, dtype=np.float32)
    n_inj = int(b.adj_inj[2].numel()) - 1
    store = b.max_nt * (b.total_n + n_inj) * 4
    # free memory just above the whole store before the first call...
    monkeypatch.setattr(torch.cuda, "mem_get_info", lambda dev=None: (int(store * 1.2), int(store * 10)))
    b.forward(keep_dw=True)
    assert b.K == b.max_nt and not b.ckpts
    # ...and, with the store now held, a fifth of it free
    monkeypatch.setattr(torch.cuda, "mem_get_info", lambda dev=None: (int(store * 0.2), int(store * 10)))
    b.forward(keep_dw=True)
    assert b.K == b.max_nt and not b.ckpts
    rng = np.random.default_rng(0)
    b.migrate([rng.standard_normal((p.nt, p.n_rec)) for p in probs])
    # an explicit budget below the store still checkpoints (each checkpoint
    # holds 6 slots' worth of state: 3 segments of 40 slots fit 0.6 of it)
    b.memory_budget = store * 6 // 10
    b.forward(keep_dw=True)
    assert b.K < b.max_nt and len(b.ckpts) >= 2
