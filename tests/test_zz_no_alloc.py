"""pysit_amd_batch_run_static (the run without a caller workspace) against
pysit_amd_batch_run, through the C ABI: FP32 imaging runs whose window
accumulator lives in the library's static scratch array give the same
images bit for bit, on two calls in a row (the second on a smaller batch,
so a stale tail of the array from the first call would show), and a batch
beyond the array is refused before anything runs."""
import ctypes

import numpy as np
import pytest

from tests.problems import Synthetic

pytestmark = pytest.mark.gpu


def _case(shape, pml, nt, seed, n_shots):
    from pysit_amd.device import DeviceBatch
    base = Synthetic(shape, pml, (0.01, 0.01), nt, 8, seed=seed)
    probs = [base] + [Synthetic(shape, pml, (0.01, 0.01), nt, 8, seed=seed + 1 + s, model=base.model,
                                dt=base.dt) for s in range(n_shots - 1)]
    b = DeviceBatch([p.geometry() for p in probs], base.model, base.deltas, base.dt, 8, dtype=np.float32,
                    graph_steps=0)
    rng = np.random.default_rng(seed)
    res = [rng.standard_normal((p.nt, p.n_rec)) for p in probs]
    return b, res


def _static_image(b, res):
    """The adjoint run of b.migrate(res), issued through the static entry
    point: same descriptor, no ic32 and no precomputed injection."""
    import torch
    from pysit_amd import _native as N
    from pysit_amd.device import _stream
    for f in b.qstate.values():
        f.zero_()
    ic = torch.zeros(b.total_n, dtype=torch.float64, device=b.dev)
    qb = b._batch(b.qstate, True, ic=ic, dw=b.dw)
    qb.ic_window = b.ic_window
    assert qb.ic_window >= 1 and not qb.ic32 and not qb.inj_pre
    at = np.ascontiguousarray(b.active_tiles)
    ar = np.zeros_like(b.active_recs)
    N.call("pysit_amd_batch_run_static", ctypes.byref(qb), 2, 0, b.max_nt, -1, 0,
           at.ctypes.data_as(ctypes.c_void_p), ar.ctypes.data_as(ctypes.c_void_p), _stream())
    torch.cuda.synchronize()
    return ic.cpu().numpy()


def test_static_run_equals_workspace_run_twice(gpu):
    cases = [((90, 110), ((10, 10), (10, 10)), 160, 4, 3),   # first call: 3 shots
             ((50, 70), ((5, 5), (4, 6)), 120, 11, 1)]       # second: one smaller shot
    for shape, pml, nt, seed, n_shots in cases:
        print(f"static run: {n_shots} shots of {shape}, nt {nt}", flush=True)
        b, res = _case(shape, pml, nt, seed, n_shots)
        b.forward(keep_dw=True)
        ref = b.migrate(res).cpu().numpy().copy()   # pysit_amd_batch_run, caller's ic32
        assert np.linalg.norm(ref) > 0
        got = _static_image(b, res)
        assert np.array_equal(got, ref)


def test_static_run_refuses_batches_beyond_the_scratch_array(gpu):
    from pysit_amd import _native as N
    b, res = _case((50, 70), ((5, 5), (4, 6)), 20, 11, 1)
    b.forward(keep_dw=True)
    b.migrate(res)
    import torch
    ic = torch.zeros(b.total_n, dtype=torch.float64, device=b.dev)
    qb = b._batch(b.qstate, True, ic=ic, dw=b.dw)
    qb.ic_window = b.ic_window
    qb.total_n = (64 << 20) // 4 + 1   # one node beyond the array: refused before any launch
    with pytest.raises(N.NativeError, match="exceeds the scratch array"):
        N.call("pysit_amd_batch_run_static", ctypes.byref(qb), 2, 0, b.max_nt, -1, 0, None, None, None)
    qb.total_n = b.total_n
    qb.ic32 = ctypes.c_void_p(ic.data_ptr()).value
    with pytest.raises(N.NativeError, match="takes no workspace"):
        N.call("pysit_amd_batch_run_static", ctypes.byref(qb), 2, 0, b.max_nt, -1, 0, None, None, None)
