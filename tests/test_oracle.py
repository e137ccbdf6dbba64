"""Pins of the CPU oracle (CPU tests).

* steps: bit-exact against the golden vectors written from PySIT's own step
  templates (tests/golden/make_golden.py) and, when it is built here, against
  oracle/_ref directly on fresh random inputs;
* taps: polynomial exactness of the centred stencils (closed form);
* forward loop: against a per-step restatement of
  pysit/modeling/temporal_modeling.py:120-158 on a tiny grid;
* migrate: with the PML off PySIT's time-reversed recursion is the exact
  discrete adjoint, so the migrated residual must equal the central finite
  difference of the misfit; checkpoint interval invariance (bit-exact);
* fold: exact transpose of 'edge' padding.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.problems import Synthetic, pml_pair

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("tag,dtype", [("f64", np.float64), ("f32", np.float32)])
@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_step2d_golden_bit_exact(tag, dtype, order):
    g = np.load(os.path.join(GOLDEN, "step2d.npz"))
    dt, dx, dz, nx, nz = g["scalars"]
    ins = [g[n] for n in ["km1_u", "k_phix", "k_phiz", "k_u", "C", "rhs"]]
    pm = [g[n] for n in ["xl", "xr", "zl", "zr"]]
    out = O.step2d(order, *ins, *pm, dt, dx, dz, int(nx), int(nz), dtype=dtype)
    for nm, a in zip(["kp1_phix", "kp1_phiz", "kp1_u"], out):
        ref = g[f"{nm}_{tag}_o{order}"]
        assert np.array_equal(a, ref), nm


@pytest.mark.parametrize("tag,dtype", [("f64", np.float64), ("f32", np.float32)])
@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_step3d_golden_bit_exact(tag, dtype, order):
    g = np.load(os.path.join(GOLDEN, "step3d.npz"))
    dt, dx, dy, dz, nx, ny, nz = g["scalars"]
    ins = [g[n] for n in ["km1_u", "k_phix", "k_phiy", "k_phiz", "k_psi", "k_u", "C", "rhs"]]
    pm = [g[n] for n in ["xl", "xr", "yl", "yr", "zl", "zr"]]
    out = O.step3d(order, *ins, *pm, dt, dx, dy, dz, int(nx), int(ny), int(nz), dtype=dtype)
    for nm, a in zip(["kp1_phix", "kp1_phiy", "kp1_phiz", "kp1_psi", "kp1_u"], out):
        ref = g[f"{nm}_{tag}_o{order}"]
        assert np.array_equal(a, ref), nm


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs the PySIT sources)")
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_steps_vs_reference_build(dtype, order):
    rng = np.random.default_rng(order)
    nx, nz = 31, 26
    f = [rng.standard_normal(nx * nz) for _ in range(6)]
    f[4] = 1 + rng.random(nx * nz)
    pm = [pml_pair(7, 33.0)[0], pml_pair(5, 80.0)[1], pml_pair(4, 20.0)[0], pml_pair(9, 41.0)[1]]
    a = O.step2d(order, *f, *pm, 7e-4, 0.02, 0.017, nx, nz, dtype=dtype)
    b = O.step2d(order, *f, *pm, 7e-4, 0.02, 0.017, nx, nz, dtype=dtype, use_ref=True)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    nx, ny, nz = 10, 12, 11
    n = nx * ny * nz
    f = [rng.standard_normal(n) for _ in range(8)]
    f[6] = 1 + rng.random(n)
    pm = [pml_pair(3, 33.0)[0], pml_pair(2, 80.0)[1], pml_pair(0, 20.0)[0], pml_pair(4, 41.0)[1],
          pml_pair(2, 50.0)[0], pml_pair(3, 10.0)[1]]
    a = O.step3d(order, *f, *pm, 7e-4, 0.02, 0.018, 0.017, nx, ny, nz, dtype=dtype)
    b = O.step3d(order, *f, *pm, 7e-4, 0.02, 0.018, 0.017, nx, ny, nz, dtype=dtype, use_ref=True)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_taps_exact_on_polynomials(order):
    """Without PML, C = 1, u_{k-1} = 0, rhs = 0: u_{k+1} = dt^2 lap(p) + 2p at
    nodes at least r away from the edges, for p of degree <= order."""
    nx, nz = 24, 22
    dx, dz, dt = 0.5, 0.25, 0.1
    x = np.arange(nx)[:, None] * dx
    z = np.arange(nz)[None, :] * dz
    deg = order
    p = 0.3 * x ** deg - 0.7 * z ** deg + x * z + 1.0
    lap = 0.3 * deg * (deg - 1) * x ** (deg - 2) - 0.7 * deg * (deg - 1) * z ** (deg - 2)
    zeros = np.zeros(nx * nz)
    e = np.zeros(0)
    _, _, up = O.step2d(order, zeros, zeros, zeros, p.reshape(-1), np.ones(nx * nz), zeros,
                        e, e, e, e, dt, dx, dz, nx, nz)
    r = order // 2
    up = up.reshape(nx, nz)
    want = dt * dt * lap + 2 * p
    sl = (slice(r, nx - r), slice(r, nz - r))
    assert np.allclose(up[sl], want[sl], rtol=1e-10, atol=1e-8)


def _python_forward(prob, nsteps):
    """Per-step restatement of forward_model's loop on the oracle step."""
    nx, nz = prob.shape
    N = nx * nz
    (xl, xr), (zl, zr) = prob.kpml
    sx, sz = prob.msig
    S1 = (sx[:, None] + sz[None, :]).reshape(-1)
    S2 = (sx[:, None] * sz[None, :]).reshape(-1)
    us = [np.zeros(N) for _ in range(3)]  # kp1, k, km1
    aux = [[np.zeros(N), np.zeros(N)] for _ in range(3)]
    sim, dws = [], []
    idx, chan, w = prob.src
    ridx, rchan, rw = prob.rec
    for k in range(nsteps):
        uk = us[1]
        d = np.zeros(prob.n_rec)
        for e in range(ridx.size):
            d[rchan[e]] += rw[e] * uk[ridx[e]]
        sim.append(d)
        rhs = np.zeros(N)
        for e in range(idx.size):
            rhs[idx[e]] += w[e] * prob.amp[k, chan[e]]
        px, pz, up = O.step2d(prob.order, us[2], aux[1][0], aux[1][1], us[1], prob.C, rhs, xl, xr, zl, zr,
                              prob.dt, prob.deltas[0], prob.deltas[1], nx, nz)
        us[0][:] = up
        aux[0][0][:] = px
        aux[0][1][:] = pz
        dws.append((us[0] - 2 * us[1] + us[2]) / prob.dt ** 2 + (S1 * (us[0] - us[2]) / (2.0 * prob.dt) + S2 * us[1]))
        if k == nsteps - 1:
            break
        us[-1] *= 0
        aux[-1][0] *= 0
        aux[-1][1] *= 0
        us.insert(0, us.pop(-1))
        aux.insert(0, aux.pop(-1))
    return np.array(sim), np.array(dws)


def test_forward_loop_matches_reference_loop():
    s = Synthetic((24, 21), ((4, 3), (3, 5)), (0.01, 0.012), 25, 4, seed=1, n_src=2, n_rec=4)
    prob = s.oracle_problem()
    out = prob.forward(simdata=True, dwaveop=True)
    sim, dws = _python_forward(prob, prob.nsteps)
    assert np.array_equal(out["simdata"], sim)
    assert np.array_equal(out["dWaveOp"], dws)


def _misfit_and_residual(prob, d_obs):
    sim = prob.forward()["simdata"]
    r = d_obs - sim
    return 0.5 * prob.dt * float(np.sum(r * r)), r


@pytest.mark.parametrize("order", [2, 8])
def test_migrate_is_exact_gradient_without_pml(order):
    """Central finite differences of J(m) = 1/2 dt sum ||S u - d||^2 against
    the migrated residual (PML off: the reverse-time recursion is the exact
    adjoint).  S carries prod(deltas) (here 0.5) and S* does not, so the
    directional derivative is prod(deltas) * <-migrate(r), dm>."""
    s = Synthetic((40, 36), ((0, 0), (0, 0)), (0.01, 0.01), 160, order, seed=2, n_src=1, n_rec=9)
    truth = s.oracle_problem()
    rng = np.random.default_rng(3)
    Ct = s.Cpad.copy()
    Ct[12:28, 14:24] *= 1.15
    truth.C = Ct.reshape(-1)
    d_obs = truth.forward()["simdata"]
    prob = s.oracle_problem()
    m0 = prob.C ** -2.0
    _, r = _misfit_and_residual(prob, d_obs)
    g = -prob.migrate(r, fold=False)
    dm = rng.standard_normal(m0.size) * m0 * 1e-2
    eps = 1e-3
    vals = []
    for sgn in (1, -1):
        p2 = s.oracle_problem()
        p2.C = (m0 + sgn * eps * dm) ** -0.5
        vals.append(_misfit_and_residual(p2, d_obs)[0])
    fd = (vals[0] - vals[1]) / (2 * eps)
    an = 0.5 * float(np.dot(g, dm))
    assert abs(fd - an) <= 1e-6 * abs(fd), (fd, an)


def test_migrate_checkpoint_interval_invariance():
    s = Synthetic((30, 27), ((5, 4), (4, 5)), (0.01, 0.01), 47, 8, seed=4)
    prob = s.oracle_problem()
    r = np.random.default_rng(5).standard_normal((prob.nsteps, prob.n_rec))
    a = prob.migrate(r, fold=False, ckpt_interval=0)
    for K in (1, 5, 13, 46):
        b = prob.migrate(r, fold=False, ckpt_interval=K)
        assert np.array_equal(a, b), K


@pytest.mark.parametrize("shape,pads", [((30, 25), ((5, 4), (3, 6))),
                                        ((12, 14, 16), ((2, 3), (4, 1), (3, 3))),
                                        ((9, 8), ((0, 0), (2, 0)))])
def test_fold_is_transpose_of_edge_padding(shape, pads):
    rng = np.random.default_rng(1)
    inner = tuple(n - l - r for n, (l, r) in zip(shape, pads))
    x = rng.standard_normal(inner)
    y = rng.standard_normal(shape)
    lhs = float(np.sum(np.pad(x, pads, mode="edge") * y))
    rhs = float(np.sum(x.reshape(-1) * O.fold_padding(y.reshape(-1), shape, pads)))
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))


def test_interp_zero_fill_closed_forms():
    ts = np.arange(11) * 0.1
    data = np.stack([2 * ts + 1, -ts], axis=1)
    tn = np.array([-0.05, 0.0, 0.033, 0.5, 0.999, 1.0, 1.01])
    out = O.interp_zero_fill(ts, data, tn)
    inside = (tn >= 0) & (tn <= 1.0)
    want = np.stack([2 * tn + 1, -tn], axis=1) * inside[:, None]
    assert np.allclose(out, want, atol=1e-14)


def test_interp_zero_fill_matches_scipy():
    """The oracle's restatement of the receivers' interpolation equals
    scipy's interp1d(kind='linear', bounds_error=False, fill_value=0) -- the
    call receivers.py:75-105 makes -- on irregular data, a finer and a
    coarser time axis and points outside the range."""
    from scipy.interpolate import interp1d
    rng = np.random.default_rng(12)
    ts = np.cumsum(rng.uniform(0.5, 1.5, 200)) * 1e-3
    data = rng.standard_normal((200, 7))
    for tn in (np.linspace(-0.01, ts[-1] + 0.01, 731), np.linspace(ts[0], ts[-1], 57),
               np.sort(rng.uniform(-0.02, ts[-1] + 0.02, 300))):
        want = interp1d(ts, data, axis=0, kind="linear", bounds_error=False, fill_value=0.0)(tn)
        got = O.interp_zero_fill(ts, data, tn)
        assert np.max(np.abs(got - want)) <= 4e-16 * np.max(np.abs(want))


@pytest.mark.parametrize("tag", ["c1", "c3d", "c1o8", "c3do8"])
def test_oracle_pml_composition_matches_reference_python(tag):
    """With the PML on, the oracle's forward loop (traces, dWaveOp with the
    mesh profile and, in 3D, the psi term), its migrate with the fold, and
    the TemporalLeastSquares objective and gradient built on them (observed
    data from the oracle in the true model, interpolated onto the start
    model's time axis) against PySIT's own Python path
    (tests/golden/pml_golden.npz, written by tests/golden/make_pml_golden.py)."""
    from tests.problems import golden_operand, oracle_from_solver, pml_golden_case, rel_l2
    g = np.load(os.path.join(GOLDEN, "pml_golden.npz"))
    m, shots, solver, C, C0 = pml_golden_case(tag)
    solver.model_parameters = solver.ModelParameters(m, {"C": C})
    ts_true = solver.ts()
    d_true = oracle_from_solver(solver, shots[0]).forward()["simdata"]
    solver.model_parameters = solver.ModelParameters(m, {"C": C0})
    assert solver.nsteps == int(g[f"{tag}_nsteps"]) and solver.dt == float(g[f"{tag}_dt"])
    op = oracle_from_solver(solver, shots[0])
    fw = op.forward(dwaveop=True)
    every = int(g[f"{tag}_every"])
    e = {"traces": rel_l2(fw["simdata"][::every], g[f"{tag}_traces"])}
    e["dWaveOp"] = max(rel_l2(fw["dWaveOp"][k], g[f"{tag}_dwaveop"][i]) for i, k in enumerate(g[f"{tag}_dw_steps"]))
    r = golden_operand(*fw["simdata"].shape)
    e["migrate"] = rel_l2(op.migrate(r), g[f"{tag}_migrate"])
    res = O.interp_zero_fill(ts_true, d_true, solver.ts()) - fw["simdata"]
    e["gradient"] = rel_l2(-op.migrate(res), g[f"{tag}_gradient"])
    f = O.misfit(res, solver.dt)
    e["objective"] = abs(f - float(g[f"{tag}_objective"])) / float(g[f"{tag}_objective"])
    print(tag, {k: f"{v:.1e}" for k, v in e.items()})
    assert all(v < 1e-10 for v in e.values()), e
