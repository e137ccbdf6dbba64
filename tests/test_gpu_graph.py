"""The time loop as replayed hipGraphs (GPU tests, through the C ABI's plan
entry points): every result equals the one-launch-per-step loop bit for bit,
across graph replays with other device work between them."""
import numpy as np
import pytest

from tests.problems import Synthetic

pytestmark = pytest.mark.gpu


def _uneven(dim=2):
    if dim == 3:
        return [Synthetic((20, 22, 70), ((4, 4), (3, 5), (4, 3)), (0.01, 0.011, 0.012), 90, 8, seed=5, n_rec=5)]
    base = Synthetic((80, 60), ((6, 6), (5, 5)), (0.01, 0.01), 100, 8, seed=3)
    probs = []
    for s, (w0, w1, nt) in enumerate([(0, 50, 100), (20, 67, 73), (5, 40, 88)]):
        window = [(6 - w0, w0, w1 - 1), (5, 0, base.model.shape[1] - 1)]
        probs.append(Synthetic((w1 - w0 + 12, 60), ((6, 6), (5, 5)), (0.01, 0.01), nt, 8, seed=20 + s,
                               window=window, model=base.model, dt=base.dt))
    return probs


def _batch(probs, dtype, graph_steps, **kw):
    from pysit_amd.device import DeviceBatch
    p0 = probs[0]
    return DeviceBatch([p.geometry() for p in probs], p0.model, p0.deltas, p0.dt, p0.order, dtype=dtype,
                       graph_steps=graph_steps, **kw)


def _gradient(b, res, segment=None):
    tr = [t.cpu().numpy().copy() for t in b.forward(keep_dw=True, segment=segment)]
    b.migrate(res)
    img = [b.image(i).cpu().numpy().copy() for i in range(b.n_shots)]
    return tr, img


def _eager_work(gpu):
    import torch
    x = torch.randn(512, 512, device=gpu)
    y = (x @ x.T).sum()
    torch.cuda.synchronize()
    return float(y)


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_graph_loop_equals_step_launches_2d_uneven(gpu, dtype):
    probs = _uneven()
    rng = np.random.default_rng(9)
    res = [rng.standard_normal((p.nt, p.n_rec)) for p in probs]
    ref = _gradient(_batch(probs, dtype, 0), res)
    b = _batch(probs, dtype, 8)
    for rep in range(3):
        out = _gradient(b, res)
        assert _same(out[0], ref[0]) and _same(out[1], ref[1]), rep
        _eager_work(gpu)
    st = b.plan.stats()
    print(st, flush=True)
    assert st["graph_launches"] > 0 and st["captures"] <= st["graph_launches"]
    assert st["direct_steps"] > 0  # the chunks where a shot starts


def test_graph_loop_checkpointed_and_windowed(gpu):
    # FP32 window of 8 slots with 16-step graphs (flushes inside the graph),
    # checkpoint segments of 40 slots (cursor store base per segment)
    p = Synthetic((90, 110), ((10, 10), (10, 10)), (0.01, 0.01), 300, 8, seed=4)
    rng = np.random.default_rng(7)
    res = [rng.standard_normal((300, p.n_rec))]
    for seg in (None, 40):
        ref = _gradient(_batch([p], np.float32, 0, ic_window=8), res, segment=seg)
        b = _batch([p], np.float32, 16, ic_window=8)
        for rep in range(2):
            out = _gradient(b, res, segment=seg)
            assert _same(out[0], ref[0]) and _same(out[1], ref[1]), (seg, rep)
            _eager_work(gpu)
        assert b.plan.stats()["graph_launches"] > 0


def test_graph_loop_born(gpu):
    p = Synthetic((60, 75), ((6, 7), (5, 8)), (0.01, 0.012), 90, 8, seed=2)
    m1 = np.random.default_rng(4).standard_normal(p.model.shape)
    outs = []
    for g in (0, 8):
        b = _batch([p], np.float64, g)
        b.forward(keep_dw=True, segment=p.nt)
        outs.append([t.cpu().numpy().copy() for t in b.born(m1)])
    assert _same(outs[0], outs[1])


def test_graph_loop_3d(gpu):
    probs = _uneven(3)
    rng = np.random.default_rng(3)
    res = [rng.standard_normal((p.nt, p.n_rec)) for p in probs]
    ref = _gradient(_batch(probs, np.float32, 0), res)
    b = _batch(probs, np.float32, 8)
    for rep in range(2):
        out = _gradient(b, res)
        assert _same(out[0], ref[0]) and _same(out[1], ref[1]), rep
        _eager_work(gpu)
    assert b.plan.stats()["graph_launches"] > 0


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_precomputed_adjoint_sources_equal_the_gathered_ones(gpu, dtype):
    """The adjoint sources evaluated for the whole run before its first step
    give the same images and adjoint fields, bit for bit, as gathering them
    in every step: uneven shots, a checkpointed run (one buffer of K slots
    reused per segment) and graph launches."""
    probs = _uneven()
    rng = np.random.default_rng(4)
    res = [rng.standard_normal((p.nt, p.n_rec)) for p in probs]
    for gs, seg in ((0, None), (8, None), (0, 30)):
        out = []
        for pre in (False, True):
            b = _batch(probs, dtype, gs)
            b.precompute_injection = pre
            out.append(_gradient(b, res, segment=seg))
            if pre:
                assert b._inj_pre is not None
            else:
                assert getattr(b, "_inj_pre", None) is None
        assert _same(out[0][0], out[1][0]) and _same(out[0][1], out[1][1]), (gs, seg)
    b = _batch(probs, dtype, 0)
    b.forward(keep_dw=True)
    q = []
    for pre in (False, True):
        b.precompute_injection = pre
        b.migrate(res, store_field=True)
        q.append(b.adjoint_store.cpu().numpy().copy())
    assert np.array_equal(q[0], q[1])


def test_precomputed_adjoint_sources_3d(gpu):
    probs = _uneven(3)
    rng = np.random.default_rng(6)
    res = [rng.standard_normal((p.nt, p.n_rec)) for p in probs]
    out = []
    for pre in (False, True):
        b = _batch(probs, np.float32, 0)
        b.precompute_injection = pre
        out.append(_gradient(b, res))
    assert _same(out[0][0], out[1][0]) and _same(out[0][1], out[1][1])
