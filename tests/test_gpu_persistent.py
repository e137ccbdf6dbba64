"""The 2D time loop as one persistent cooperative launch per run (GPU tests,
through pysit_amd_batch_run_persistent): every result equals the
one-launch-per-step loop bit for bit."""
import numpy as np
import pytest

from tests.problems import Synthetic
from tests.test_gpu_graph import _batch, _eager_work, _gradient, _same, _uneven

pytestmark = pytest.mark.gpu


def _persistent(probs, dtype, **kw):
    b = _batch(probs, dtype, 0, **kw)
    b.persistent = True
    return b


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("segment", [None, 30])
def test_persistent_loop_equals_step_launches_uneven(gpu, dtype, segment):
    """Uneven shots (their active tile and receiver counts change from step
    to step), store-all and checkpointed (recompute runs of mode 1), FP64
    imaging and the FP32 window's flushes inside the launch."""
    probs = _uneven()
    rng = np.random.default_rng(9)
    res = [rng.standard_normal((p.nt, p.n_rec)) for p in probs]
    print("persistent uneven", dtype.__name__, segment, flush=True)
    ref = _gradient(_batch(probs, dtype, 0), res, segment=segment)
    b = _persistent(probs, dtype)
    for rep in range(2):
        out = _gradient(b, res, segment=segment)
        assert _same(out[0], ref[0]) and _same(out[1], ref[1]), rep
        _eager_work(gpu)


@pytest.mark.parametrize("window", [7, 32])
def test_persistent_loop_fp32_windows(gpu, window):
    """FP32 windows of 7 slots (the last one partial) and 32 slots, with
    checkpointed segments of 50 slots: the flushes the launch makes between
    its steps land where the step loop's k_ic_flush launches do."""
    p = Synthetic((90, 110), ((10, 10), (10, 10)), (0.01, 0.01), 300, 8, seed=4)
    rng = np.random.default_rng(7)
    res = [rng.standard_normal((300, p.n_rec))]
    for seg in (None, 50):
        print("persistent window", window, seg, flush=True)
        ref = _gradient(_batch([p], np.float32, 0, ic_window=window), res, segment=seg)
        out = _gradient(_persistent([p], np.float32, ic_window=window), res, segment=seg)
        assert _same(out[0], ref[0]) and _same(out[1], ref[1]), seg


@pytest.mark.parametrize("order", [2, 4, 6])
def test_persistent_loop_orders(gpu, order):
    p = Synthetic((60, 75), ((6, 7), (5, 8)), (0.01, 0.012), 90, order, seed=2)
    rng = np.random.default_rng(1)
    res = [rng.standard_normal((p.nt, p.n_rec))]
    print("persistent order", order, flush=True)
    for dtype in (np.float64, np.float32):
        ref = _gradient(_batch([p], dtype, 8), res)
        out = _gradient(_persistent([p], dtype), res)
        assert _same(out[0], ref[0]) and _same(out[1], ref[1]), dtype


def test_persistent_loop_leaves_born_to_step_launches(gpu):
    """Born runs (mode 3) and the second-order terms are not in the
    persistent kernel: a persistent batch runs them as step launches."""
    p = Synthetic((60, 75), ((6, 7), (5, 8)), (0.01, 0.012), 90, 8, seed=2)
    m1 = np.random.default_rng(4).standard_normal(p.model.shape)
    outs = []
    for persistent in (False, True):
        b = _batch([p], np.float64, 0)
        b.persistent = persistent
        b.forward(keep_dw=True, segment=p.nt)
        outs.append([t.cpu().numpy().copy() for t in b.born(m1)])
    assert _same(outs[0], outs[1])


def test_persistent_loop_reference_case(gpu):
    """PySIT's own horizontal-reflector case (91 x 71, order 4, FP64, 3 s):
    traces and gradient equal the graph loop's bit for bit."""
    import pysit_amd as pa
    from tests.problems import c1_setup
    m, shots, solver, model = c1_setup(order=4, tf=3.0)
    solver.model_parameters = model
    out = []
    for persistent in (False, True):
        tm = pa.TemporalModeling(solver)
        b = tm.batch(shots)
        b.persistent = persistent
        tr = [t.cpu().numpy().copy() for t in b.forward(keep_dw=True)]
        rng = np.random.default_rng(2)
        b.migrate([rng.standard_normal(t.shape) for t in tr])
        out.append((tr, b.ic.cpu().numpy().copy()))
    assert _same(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_persistent_loop_barrier_time_out_is_an_error(gpu, monkeypatch):
    """A barrier wait beyond the spin limit ends every block and the call
    reports it; the host forces it with a limit of one poll (every block's
    wait is bounded, so the launch still ends normally).  The batch then
    runs again at the default limit and gives the step loop's results."""
    from pysit_amd import _native as N
    probs = _uneven()
    rng = np.random.default_rng(9)
    res = [rng.standard_normal((p.nt, p.n_rec)) for p in probs]
    ref = _gradient(_batch(probs, np.float64, 0), res)
    b = _persistent(probs, np.float64)
    monkeypatch.setenv("PYSIT_AMD_PERSIST_SPIN_LIMIT", "1")
    print("persistent loop with a one-poll barrier limit", flush=True)
    with pytest.raises(N.NativeError, match="timed out"):
        b.forward(keep_dw=True)
    monkeypatch.delenv("PYSIT_AMD_PERSIST_SPIN_LIMIT")
    out = _gradient(b, res)
    assert _same(out[0], ref[0]) and _same(out[1], ref[1])
