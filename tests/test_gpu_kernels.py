"""HIP kernels against the CPU oracle (GPU tests, called through the C ABI)."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from tests.problems import Synthetic, pml_pair, rel_l2

pytestmark = pytest.mark.gpu


def _dev_arrays(gpu, arrays, dtype):
    import torch
    t = [torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to(gpu) for a in arrays]
    return t


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() else ctypes.c_void_p(0)


def _random_state_2d(rng, nx, nz):
    f = [rng.standard_normal(nx * nz) for _ in range(6)]
    f[4] = 1.0 + 0.5 * rng.random(nx * nz)  # C
    return f


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("order", [2, 4, 6, 8])
@pytest.mark.parametrize("shape", [(23, 19), (70, 133)])
def test_step2d_matches_oracle(gpu, dtype, order, shape):
    import torch
    from pysit_amd import _native as N
    rng = np.random.default_rng(order + shape[0])
    nx, nz = shape
    f = _random_state_2d(rng, nx, nz)
    xl, _ = pml_pair(5, 40.0)
    _, xr = pml_pair(6, 55.0)
    zl, _ = pml_pair(4, 70.0)
    _, zr = pml_pair(3, 30.0)
    dt, dx, dz = 1e-3, 0.011, 0.013
    ref = O.step2d(order, *f, xl, xr, zl, zr, dt, dx, dz, nx, nz, dtype=dtype)
    d = _dev_arrays(gpu, f + [xl, xr, zl, zr], dtype)
    outs = [torch.zeros(nx * nz, dtype=d[0].dtype, device=gpu) for _ in range(3)]
    name = "pysit_amd_cda_step2d_" + ("f64" if dtype == np.float64 else "f32")
    args = [order] + [_p(t) for t in d[:6]]
    for t in d[6:]:
        args += [_p(t), t.numel()]
    args += [dt, dx, dz, nx, nz] + [_p(t) for t in outs] + [None]
    N.call(name, *args)
    torch.cuda.synchronize()
    tol = 1e-12 if dtype == np.float64 else 2e-6
    for o, r in zip(outs, ref):
        got = o.cpu().numpy().astype(np.float64)
        scale = max(1.0, np.abs(r).max())
        assert np.abs(got - r).max() <= tol * scale * 10, (name, order, np.abs(got - r).max())


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_step3d_matches_oracle(gpu, dtype, order):
    # nx = 37: two 16-plane x tiles and a partial third, so the step crosses
    # x-tile boundaries at every order
    import torch
    from pysit_amd import _native as N
    rng = np.random.default_rng(7 + order)
    nx, ny, nz = 37, 17, 71
    n = nx * ny * nz
    f = [rng.standard_normal(n) for _ in range(8)]
    f[6] = 1.0 + 0.5 * rng.random(n)
    pmls = [pml_pair(3, 40.0)[0], pml_pair(4, 50.0)[1], pml_pair(2, 60.0)[0], pml_pair(3, 45.0)[1],
            pml_pair(5, 35.0)[0], pml_pair(2, 65.0)[1]]
    dt, dx, dy, dz = 1e-3, 0.011, 0.012, 0.013
    ref = O.step3d(order, *f, *pmls, dt, dx, dy, dz, nx, ny, nz, dtype=dtype)
    d = _dev_arrays(gpu, f + pmls, dtype)
    outs = [torch.zeros(n, dtype=d[0].dtype, device=gpu) for _ in range(5)]
    name = "pysit_amd_cda_step3d_" + ("f64" if dtype == np.float64 else "f32")
    args = [order] + [_p(t) for t in d[:8]]
    for t in d[8:]:
        args += [_p(t), t.numel()]
    args += [dt, dx, dy, dz, nx, ny, nz] + [_p(t) for t in outs] + [None]
    N.call(name, *args)
    torch.cuda.synchronize()
    tol = 1e-12 if dtype == np.float64 else 2e-6
    for o, r in zip(outs, ref):
        got = o.cpu().numpy().astype(np.float64)
        scale = max(1.0, np.abs(r).max())
        assert np.abs(got - r).max() <= tol * scale * 10, (name, order, np.abs(got - r).max())


def _batch(probs, dtype, **kw):
    from pysit_amd.device import DeviceBatch
    p0 = probs[0]
    return DeviceBatch([p.geometry() for p in probs], p0.model, p0.deltas, p0.dt, p0.order,
                       dtype=dtype, **kw)


@pytest.mark.parametrize("order", [4, 8])
def test_batch_forward_and_migrate_fp64_2d(gpu, order):
    probs = [Synthetic((60, 75), ((6, 7), (5, 8)), (0.01, 0.012), 90, order, seed=s, n_src=1 + s)
             for s in range(3)]
    for p in probs[1:]:
        p.model = probs[0].model
        p.Cpad = probs[0].Cpad
        p.dt = probs[0].dt
    b = _batch(probs, np.float64)
    traces = b.forward(keep_dw=True)
    rng = np.random.default_rng(5)
    res = [rng.standard_normal((p.nt, p.n_rec)) for p in probs]
    ic = b.migrate(res).cpu().numpy()
    for i, p in enumerate(probs):
        op = p.oracle_problem()
        # sources differ per shot: rebuild the oracle amplitude set for that shot
        ref = op.forward(simdata=True)["simdata"]
        assert rel_l2(traces[i].cpu().numpy(), ref) < 1e-11
        img_ref = op.migrate(res[i], fold=False)
        img = b.image(i).cpu().numpy().reshape(-1) * (-p.dt)
        assert rel_l2(img, img_ref) < 1e-10


def test_batch_checkpointed_equals_store_all(gpu):
    p = Synthetic((50, 70), ((5, 5), (4, 6)), (0.01, 0.01), 61, 8, seed=11)
    rng = np.random.default_rng(1)
    res = [rng.standard_normal((p.nt, p.n_rec))]
    b1 = _batch([p], np.float64)
    b1.forward(keep_dw=True)
    ic1 = b1.migrate(res).cpu().numpy()
    b2 = _batch([p], np.float64)
    b2.forward(keep_dw=True, segment=7)
    assert len(b2.ckpts) == 9
    ic2 = b2.migrate(res).cpu().numpy()
    assert np.array_equal(ic1, ic2)
    # the forward stored the last segment's dWaveOp; a second migrate finds
    # the first segment's in the store and recomputes the last one
    ic3 = b2.migrate(res).cpu().numpy()
    assert np.array_equal(ic1, ic3)


@pytest.mark.parametrize("budget_kib", [2150, 1920])
def test_nested_checkpoint_levels_equal_store_all(gpu, budget_kib):
    # memory budgets below what one level of checkpoints needs (240 steps of
    # a 50 x 70 FP64 shot): the schedule nests two or three levels (replays
    # with sub-checkpoints) and gives the same bits as storing every step
    p = Synthetic((50, 70), ((5, 5), (4, 6)), (0.01, 0.01), 240, 8, seed=11)
    rng = np.random.default_rng(1)
    res = [rng.standard_normal((p.nt, p.n_rec))]
    b1 = _batch([p], np.float64)
    b1.forward(keep_dw=True)
    ic1 = b1.migrate(res).cpu().numpy()
    b2 = _batch([p], np.float64, memory_budget=budget_kib * 1024)
    b2.forward(keep_dw=True)
    print("checkpoint levels", b2.levels, "top checkpoints", len(b2.ckpts))
    assert len(b2.levels) >= 2
    ic2 = b2.migrate(res).cpu().numpy()
    assert np.array_equal(ic1, ic2)


@pytest.mark.parametrize("snaps", [0, 2, 5])
def test_binomial_schedule_equals_store_all(gpu, snaps):
    """The binomial (Revolve) schedule reverses the run from rest with a few
    state snapshots, recomputing each step's dWaveOp: the same bits as
    storing every step, uneven shots included (each shot's level at a
    snapshot slot), and exactly the forward advances binomial_cost counts;
    the pseudo-Hessian diagonal sums the recomputed dWaveOp's squares.""" 
    from pysit_amd.device import DeviceBatch, binomial_cost
    base = Synthetic((60, 50), ((6, 6), (5, 5)), (0.01, 0.01), 90, 8, seed=3)
    probs = []
    for s, (w0, w1, nt) in enumerate([(0, 40, 90), (10, 47, 61), (5, 30, 77)]):
        window = [(6 - w0, w0, w1 - 1), (5, 0, base.model.shape[1] - 1)]
        probs.append(Synthetic((w1 - w0 + 12, 50), ((6, 6), (5, 5)), (0.01, 0.01), nt, 8, seed=30 + s,
                               window=window, model=base.model, dt=base.dt))
    rng = np.random.default_rng(4)
    res = [rng.standard_normal((p.nt, p.n_rec)) for p in probs]
    out = []
    for schedule in ("auto", "binomial"):
        b = DeviceBatch([p.geometry() for p in probs], base.model, base.deltas, base.dt, 8, dtype=np.float64)
        b.schedule = schedule
        b.max_snaps = snaps
        tr = [t.cpu().numpy().copy() for t in b.forward(keep_dw=True)]
        b.migrate(res, dw_squares=True)
        out.append((tr, b.ic.cpu().numpy().copy(), b.dw_sq.cpu().numpy().copy()))
        if schedule == "binomial":
            assert b.snaps == snaps and b.binomial_advances == binomial_cost(b.max_nt, snaps)
    assert all(np.array_equal(a, c) for a, c in zip(out[0][0], out[1][0]))
    assert np.array_equal(out[0][1], out[1][1])
    # the squares are summed slot by slot in reverse order there
    assert np.linalg.norm(out[0][2] - out[1][2]) <= 1e-13 * np.linalg.norm(out[0][2])


def test_binomial_schedule_is_the_fallback_below_the_nested_one(gpu):
    """A budget below every nested schedule (four levels) but above a few
    state snapshots plans the binomial schedule instead of raising; FP32."""
    from pysit_amd.device import DeviceBatch
    p = Synthetic((50, 70), ((5, 5), (4, 6)), (0.01, 0.01), 240, 8, seed=11)
    res = [np.random.default_rng(2).standard_normal((p.nt, p.n_rec))]
    b1 = _batch([p], np.float32)
    b1.forward(keep_dw=True)
    ic1 = b1.migrate(res).cpu().numpy()
    probe = _batch([p], np.float32)
    state, ckpt = probe.state_bytes(), probe.ckpt_bytes()
    n_inj = int(probe.adj_inj[2].numel()) - 1
    per_slot = (probe.total_n + n_inj) * 4
    budget = 2 * state + 12 * probe.total_n + per_slot + 3 * ckpt + 100
    b2 = _batch([p], np.float32, memory_budget=budget)
    b2.forward(keep_dw=True)
    assert b2.snaps == 3 and b2.ckpts == []
    ic2 = b2.migrate(res).cpu().numpy()
    # FP32 windows follow the slots, not the one-step adjoint runs: the same bits
    assert np.array_equal(ic2, ic1)


@pytest.mark.parametrize("budget_kib", [6000, 3000])
def test_retimed_batch_equals_a_fresh_one(gpu, budget_kib):
    """DeviceBatch.retime keeps the tables and takes a new time axis (new
    step counts and amplitudes at another dt); with segment checkpoints
    (6000 KiB) or binomial snapshots (3000 KiB) and uneven shots the traces
    and images equal a fresh batch's bit for bit, before and after going
    back to the first axis."""
    def shots(nts, dt=None):
        out = []
        for k, nt in enumerate(nts):
            out.append(Synthetic((50, 70), ((5, 5), (4, 6)), (0.01, 0.01), nt, 8, seed=11 + k,
                                 model=shots.model, dt=dt))
        return out
    shots.model = Synthetic((50, 70), ((5, 5), (4, 6)), (0.01, 0.01), 10, 8, seed=11).model
    a = shots((240, 200))
    c = shots((220, 180), dt=a[0].dt * 1.09)
    rng = np.random.default_rng(3)
    res_a = [rng.standard_normal((p.nt, p.n_rec)) for p in a]
    res_c = [rng.standard_normal((p.nt, p.n_rec)) for p in c]
    budget = budget_kib * 1024

    def run(b, probs, res):
        tr = [t.cpu().numpy().copy() for t in b.forward(keep_dw=True)]
        return tr, b.migrate(res).cpu().numpy().copy(), list(b.levels)

    b = _batch(a, np.float64, memory_budget=budget)
    first = run(b, a, res_a)
    b.retime([p.geometry() for p in c], c[0].dt)
    got = run(b, c, res_c)
    want = run(_batch(c, np.float64, memory_budget=budget), c, res_c)
    print("levels", first[2], got[2], want[2], "snaps", b.snaps)
    assert got[2] == want[2] and (b.snaps is None) == (budget_kib == 6000)
    assert all(np.array_equal(x, y) for x, y in zip(got[0], want[0]))
    assert np.array_equal(got[1], want[1])
    b.retime([p.geometry() for p in a], a[0].dt)
    again = run(b, a, res_a)
    assert all(np.array_equal(x, y) for x, y in zip(again[0], first[0]))
    assert np.array_equal(again[1], first[1])
    with pytest.raises(ValueError):
        b.retime([p.geometry() for p in a[::-1]], a[0].dt)


def test_batch_uneven_shots_checkpointed_equal_store_all(gpu):
    # uneven shots sit at different step parities at a checkpoint slot: each
    # shot's current level is saved and restored
    base = Synthetic((80, 60), ((6, 6), (5, 5)), (0.01, 0.01), 100, 8, seed=3)
    probs = []
    for s, (w0, w1, nt) in enumerate([(0, 50, 100), (20, 67, 73), (5, 40, 88)]):
        window = [(6 - w0, w0, w1 - 1), (5, 0, base.model.shape[1] - 1)]
        probs.append(Synthetic((w1 - w0 + 12, 60), ((6, 6), (5, 5)), (0.01, 0.01), nt, 8, seed=20 + s,
                               window=window, model=base.model, dt=base.dt))
    from pysit_amd.device import DeviceBatch
    rng = np.random.default_rng(9)
    res = [rng.standard_normal((p.nt, p.n_rec)) for p in probs]
    out = []
    for seg in (None, 7):
        b = DeviceBatch([p.geometry() for p in probs], base.model, base.deltas, base.dt, 8, dtype=np.float64)
        tr = [t.cpu().numpy().copy() for t in b.forward(keep_dw=True, segment=seg)]
        b.migrate(res)
        out.append((tr, b.ic.cpu().numpy().copy()))
    assert all(np.array_equal(a, c) for a, c in zip(out[0][0], out[1][0]))
    assert np.array_equal(out[0][1], out[1][1])


def test_batch_uneven_shots_2d(gpu):
    base = Synthetic((80, 60), ((6, 6), (5, 5)), (0.01, 0.01), 100, 8, seed=3)
    probs = []
    for s, (w0, w1, nt) in enumerate([(0, 50, 100), (20, 67, 73), (5, 40, 88)]):
        ncols = w1 - w0
        window = [(6 - w0, w0, w1 - 1), (5, 0, base.model.shape[1] - 1)]
        p = Synthetic((ncols + 12, 60), ((6, 6), (5, 5)), (0.01, 0.01), nt, 8, seed=20 + s,
                      window=window, model=base.model, dt=base.dt)
        probs.append(p)
    from pysit_amd.device import DeviceBatch
    b = DeviceBatch([p.geometry() for p in probs], base.model, base.deltas, base.dt, 8, dtype=np.float64)
    traces = b.forward(keep_dw=True)
    rng = np.random.default_rng(9)
    res = [rng.standard_normal((p.nt, p.n_rec)) for p in probs]
    b.migrate(res)
    for i, p in enumerate(probs):
        op = p.oracle_problem()
        ref = op.forward()["simdata"]
        assert rel_l2(traces[i].cpu().numpy(), ref) < 1e-11
        img_ref = op.migrate(res[i], fold=False)
        img = b.image(i).cpu().numpy().reshape(-1) * (-p.dt)
        assert rel_l2(img, img_ref) < 1e-10


def test_batch_fp32_close_to_fp64_oracle(gpu):
    p = Synthetic((90, 110), ((10, 10), (10, 10)), (0.01, 0.01), 400, 8, seed=4)
    b = _batch([p], np.float32)
    tr = b.forward(keep_dw=True)[0].cpu().numpy()
    op = p.oracle_problem()
    ref = op.forward()["simdata"]
    rng = np.random.default_rng(2)
    res = [ref - 0.9 * ref + 0.01 * rng.standard_normal(ref.shape) * np.abs(ref).max()]
    b.migrate(res)
    img = b.image(0).cpu().numpy().reshape(-1) * (-p.dt)
    img_ref = op.migrate(res[0], fold=False)
    e_tr, e_img = rel_l2(tr, ref), rel_l2(img, img_ref)
    print(f"FP32, 400 steps: traces {e_tr:.2e}, gradient {e_img:.2e}")
    assert e_tr < 1e-5 and e_img < 1e-4


def test_fp32_accuracy_over_a_long_run(gpu):
    # 2000 steps in FP32 against the FP64 oracle.  The batch steps evaluate
    # the Laplacian with difference-form taps and the update in increment
    # form: 2.3e-6 (traces) and 1.8e-6 (gradient) here, where the plain form
    # (the reference's own FP32 arithmetic) gave 1.5e-5 and 9e-6.
    p = Synthetic((90, 110), ((10, 10), (10, 10)), (0.01, 0.01), 2000, 8, seed=4)
    b = _batch([p], np.float32)
    tr = b.forward(keep_dw=True)[0].cpu().numpy()
    op = p.oracle_problem()
    ref = op.forward()["simdata"]
    rng = np.random.default_rng(2)
    res = [0.1 * ref + 0.01 * rng.standard_normal(ref.shape) * np.abs(ref).max()]
    b.migrate(res)
    img = b.image(0).cpu().numpy().reshape(-1) * (-p.dt)
    img_ref = op.migrate(res[0], fold=False)
    e_tr, e_img = rel_l2(tr, ref), rel_l2(img, img_ref)
    print(f"FP32, 2000 steps: traces {e_tr:.2e}, gradient {e_img:.2e}")
    assert e_tr < 1e-5 and e_img < 1e-5, (e_tr, e_img)


def test_ic_window_accumulator_matches_fp64_accumulation(gpu):
    # FP32 batch: imaging sums in FP32 windows of 7 slots (400 steps: the last
    # window is partial), flushed into FP64, against FP64 accumulation every
    # step; checkpointed segments of 50 slots (windows span segment ends)
    p = Synthetic((90, 110), ((10, 10), (10, 10)), (0.01, 0.01), 400, 8, seed=4)
    rng = np.random.default_rng(7)
    res = [rng.standard_normal((400, p.n_rec))]
    imgs = []
    for w, seg in ((0, None), (7, None), (7, 50), (32, None)):
        b = _batch([p], np.float32, ic_window=w)
        b.forward(keep_dw=True, segment=seg)
        b.migrate(res)
        imgs.append(b.image(0).cpu().numpy().reshape(-1).copy())
    ref = imgs[0]
    for img in imgs[1:]:
        assert rel_l2(img, ref) < 1e-5


@pytest.mark.parametrize("window", [7, 64])
def test_fp32_windows_checkpointed_equal_store_all(gpu, window):
    """The FP32 imaging windows follow the batch's slots, not the runs: the
    checkpointed sweeps (segments of 50 and 37 slots, nested levels, the
    binomial schedule) sum every window exactly as the store-all sweep, so
    their FP32 images are the store-all one bit for bit; uneven shots."""
    from pysit_amd.device import DeviceBatch
    base = Synthetic((60, 50), ((6, 6), (5, 5)), (0.01, 0.01), 190, 8, seed=3)
    probs = []
    for s, (w0, w1, nt) in enumerate([(0, 40, 190), (10, 47, 131), (5, 30, 160)]):
        window_ = [(6 - w0, w0, w1 - 1), (5, 0, base.model.shape[1] - 1)]
        probs.append(Synthetic((w1 - w0 + 12, 50), ((6, 6), (5, 5)), (0.01, 0.01), nt, 8, seed=30 + s,
                               window=window_, model=base.model, dt=base.dt))
    rng = np.random.default_rng(4)
    res = [rng.standard_normal((p.nt, p.n_rec)) for p in probs]

    def run(segment=None, schedule="auto", budget=None, graph_steps=None):
        b = DeviceBatch([p.geometry() for p in probs], base.model, base.deltas, base.dt, 8, dtype=np.float32,
                        ic_window=window, memory_budget=budget, graph_steps=graph_steps)
        b.schedule = schedule
        if schedule == "binomial":
            b.max_snaps = 3
        b.forward(keep_dw=True, segment=segment)
        return b.migrate(res).cpu().numpy().copy(), b
    ref, b0 = run()
    assert np.linalg.norm(ref) > 0
    print("FP32 windows", window, flush=True)
    for kw in ({"segment": 50}, {"segment": 37}, {"segment": 50, "graph_steps": 0}, {"schedule": "binomial"}):
        got, b = run(**kw)
        assert np.array_equal(got, ref), kw
    # nested levels: a budget below one-level checkpointing
    state, ckpt = b0.state_bytes(), b0.ckpt_bytes()
    n_inj = int(b0.adj_inj[2].numel()) - 1
    per_slot = (b0.total_n + n_inj) * 4
    got, b = run(budget=2 * state + 12 * b0.total_n + 12 * per_slot + 8 * ckpt)
    assert len(b.levels) >= 2, b.levels
    assert np.array_equal(got, ref)


def test_ic_window_accumulator_3d_matches_fp64_accumulation(gpu):
    # the same in 3D (k_step3d mode 4): windows of 7 and 64 slots, and
    # checkpointed segments of 30 slots, against FP64 accumulation every step
    p = Synthetic((24, 26, 70), ((4, 4), (3, 5), (4, 3)), (0.01, 0.011, 0.012), 130, 8, seed=8, n_rec=6)
    rng = np.random.default_rng(9)
    res = [rng.standard_normal((p.nt, p.n_rec))]
    imgs = []
    for w, seg in ((0, None), (7, None), (7, 30), (64, None)):
        b = _batch([p], np.float32, ic_window=w)
        assert b.ic_window == w
        b.forward(keep_dw=True, segment=seg)
        b.migrate(res)
        imgs.append(b.image(0).cpu().numpy().reshape(-1).copy())
    ref = imgs[0]
    assert np.linalg.norm(ref) > 0
    for img in imgs[1:]:
        assert rel_l2(img, ref) < 1e-5


@pytest.mark.parametrize("dtype,tol,order", [(np.float64, 1e-10, 4), (np.float64, 1e-10, 6), (np.float64, 1e-10, 8),
                                             (np.float32, 1e-5, 8)])
def test_batch_3d(gpu, dtype, tol, order):
    p = Synthetic((20, 22, 70), ((4, 4), (3, 5), (4, 3)), (0.01, 0.011, 0.012), 90, order, seed=5, n_rec=5)
    b = _batch([p], dtype)
    tr = b.forward(keep_dw=True)[0].cpu().numpy()
    op = p.oracle_problem()
    ref = op.forward()["simdata"]
    assert rel_l2(tr, ref) < tol
    rng = np.random.default_rng(3)
    res = [rng.standard_normal(ref.shape)]
    b.migrate(res)
    img = b.image(0).cpu().numpy().reshape(-1) * (-p.dt)
    img_ref = op.migrate(res[0], fold=False)
    assert rel_l2(img, img_ref) < tol * 10


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 1e-5)])
def test_batch_3d_with_plain_tiles(gpu, dtype, tol):
    # 64 x 36 x 140 nodes with 4-node PML: the x-y tiles in the middle (x
    # tiles of 16 planes, y tiles of 8 rows) and the z tile at 64..127 are
    # plain tiles (no PML node within r of them); the source and the
    # receivers sit in them, and the whole last wavefield is compared too
    p = Synthetic((64, 36, 140), ((4, 4), (4, 4), (4, 4)), (0.01, 0.011, 0.012), 60, 8, seed=7, n_rec=9,
                  src_center=(32, 18, 92), rec_z=96)
    b = _batch([p], dtype)
    flags = b.tiles.view(-1, 4)[:, 2].cpu().numpy()
    assert ((flags & 1) == 0).sum() >= 4  # bit0: the PML path
    tr = b.forward(keep_dw=True, store_field=True)[0].cpu().numpy()
    field = b.shot_field(b.field_store, 0, p.nt - 1).cpu().numpy().reshape(-1)
    op = p.oracle_problem()
    fw = op.forward(wavefield=True)
    ref = fw["simdata"]
    assert np.linalg.norm(ref) > 1e-3 * np.abs(p.amp).max()
    assert rel_l2(tr, ref) < tol
    assert rel_l2(field, fw["wavefield"][p.nt - 1]) < tol
    rng = np.random.default_rng(3)
    res = [rng.standard_normal(ref.shape)]
    b.migrate(res)
    img = b.image(0).cpu().numpy().reshape(-1) * (-p.dt)
    img_ref = op.migrate(res[0], fold=False)
    assert rel_l2(img, img_ref) < tol * 10


def test_layered_3d_workload_is_stable_and_fp32_follows_fp64(gpu):
    # the 3D workload (km, km/s) on 48^3 nodes + 16 PML cells, 600 steps: the
    # traces stay bounded and FP32 follows FP64
    from pysit_amd.device import DeviceBatch
    from pysit_amd.modeling import TemporalModeling
    from pysit_amd.workloads import layered_3d
    mesh, shots, solver, model = layered_3d(n=48, n_shots=1, nt=600, rec_patch=16)
    tm = TemporalModeling(solver, autotune=False)
    g = tm.geometry(shots[0])
    trs = []
    for dtype in (np.float64, np.float32):
        b = DeviceBatch([g], solver.unpadded_model(), mesh.deltas, solver.dt, 8, dtype=dtype)
        trs.append(b.forward(keep_dw=False)[0].cpu().numpy())
    t64, t32 = trs
    assert np.all(np.isfinite(t64)) and np.all(np.isfinite(t32))
    late, peak = np.abs(t64[-100:]).max(), np.abs(t64).max()
    assert late < peak
    assert rel_l2(t32, t64) < 1e-5


def test_fold_matches_oracle(gpu):
    import torch
    from pysit_amd import _native as N
    rng = np.random.default_rng(0)
    for shape, pads in [((30, 25), ((5, 4), (3, 6))), ((12, 14, 16), ((2, 3), (4, 1), (3, 3)))]:
        a = rng.standard_normal(int(np.prod(shape)))
        ref = O.fold_padding(a, shape, pads)
        inner = tuple(n - l - r for n, (l, r) in zip(shape, pads))
        t = torch.from_numpy(a).to(gpu)
        out = torch.zeros(inner, dtype=torch.float64, device=gpu)
        dim = len(shape)
        if dim == 2:
            args = (2, shape[0], 1, shape[1], pads[0][0], pads[0][1], 0, 0, pads[1][0], pads[1][1])
            ony, onz = 1, inner[1]
        else:
            args = (3, *shape, pads[0][0], pads[0][1], pads[1][0], pads[1][1], pads[2][0], pads[2][1])
            ony, onz = inner[1], inner[2]
        N.call("pysit_amd_fold", *args, ctypes.c_void_p(t.data_ptr()), 1.0,
               ctypes.c_void_p(out.data_ptr()), ony, onz, 0, None)
        torch.cuda.synchronize()
        assert np.allclose(out.cpu().numpy().reshape(-1), ref, rtol=1e-13, atol=1e-13)
