"""PySIT-API paths on the GPU against the oracle (GPU tests).

C1 is the reference's own CPU case: the 91 x 71 horizontal reflector,
order 4, T = 1.5 (1350 steps), PML 0.1 with amplitude 100, one Ricker
source and 91 receivers, FP64.
"""
import copy

import numpy as np
import pytest

from tests.problems import c1_setup, oracle_from_solver, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1():
    m, shots, solver, model = c1_setup(order=4)
    solver.model_parameters = model
    return m, shots, solver, model


def test_c1_forward_fp64_matches_oracle(gpu, c1):
    import pysit_amd as pa
    m, shots, solver, model = c1
    tm = pa.TemporalModeling(solver)
    rv = tm.forward_model(shots[0], model, return_parameters=["simdata"])
    ref = oracle_from_solver(solver, shots[0]).forward()["simdata"]
    assert rv["simdata"].shape == (1350, 91)
    assert rel_l2(rv["simdata"], ref) < 1e-11


def test_c1_migrate_fp64_matches_oracle(gpu, c1):
    import pysit_amd as pa
    m, shots, solver, model = c1
    tm = pa.TemporalModeling(solver)
    dW = tm.forward_model(shots[0], model, return_parameters=["dWaveOp"])["dWaveOp"]
    assert len(dW) == 1350
    r = np.random.default_rng(0).standard_normal((1350, 91))
    g = tm.migrate_shot(shots[0], model, r, 1, dWaveOp=dW)
    ref = oracle_from_solver(solver, shots[0]).migrate(r)
    assert g.data.shape == (91 * 71, 1)
    assert rel_l2(g.data.reshape(-1), ref) < 1e-10


def test_c1_imaging_period_and_checkpoints(gpu, c1):
    import pysit_amd as pa
    m, shots, solver, model = c1
    r = np.random.default_rng(1).standard_normal((1350, 91))
    ref = oracle_from_solver(solver, shots[0]).migrate(r, imaging_period=3)
    tm = pa.TemporalModeling(solver, memory_budget=64 * 2 ** 20)  # forces checkpoints
    g = tm.migrate_shot(shots[0], model, r, 3)
    assert rel_l2(g.data.reshape(-1), ref) < 1e-10


def test_tls_gradient_and_objective(gpu, c1):
    import pysit_amd as pa
    m, shots, solver, model = c1
    shots = copy.deepcopy(shots)
    pa.generate_seismic_data(shots, solver, model)
    C0 = model.C.copy()
    C0[:] = 1.0
    m0 = solver.ModelParameters(m, {"C": C0})
    obj = pa.TemporalLeastSquares(solver)
    aux = {"objective_value": (True, None), "residual_norm": (True, None)}
    g = obj.compute_gradient(shots, m0, aux_info=aux)
    f = obj.evaluate(shots, m0)
    assert abs(aux["objective_value"][1] - f) <= 1e-12 * abs(f)
    # oracle: residual from the oracle forward in m0, migrate, minus sign
    solver.model_parameters = m0
    op = oracle_from_solver(solver, shots[0])
    d_obs = shots[0].receivers.interpolate_data(solver.ts())
    r = d_obs - op.forward()["simdata"]
    ref = -op.migrate(r)
    assert rel_l2(g.data.reshape(-1), ref) < 1e-10
    assert abs(0.5 * solver.dt * np.sum(r * r) - f) <= 1e-10 * f


def test_born_matches_finite_difference(gpu):
    import pysit_amd as pa
    m, shots, solver, model = c1_setup(order=4, tf=0.6)
    tm = pa.TemporalModeling(solver)
    rng = np.random.default_rng(2)
    X, Z = m.mesh_coords()
    # a bump away from the fast reflector: max C (hence dt and nsteps) stays put
    dm = (0.05 * np.exp(-((X - 0.5) ** 2 + (Z - 0.3) ** 2) / 0.01) * (Z < 0.36)).reshape(-1, 1)
    m1 = model.perturbation(data=dm)
    d1 = tm.linear_forward_model(shots[0], model, m1, return_parameters=["simdata"])["simdata"]
    eps = 1e-4
    lin = model.linearize()
    mp = solver.ModelParameters(m, linear_inputs=lin + eps * dm)
    mm = solver.ModelParameters(m, linear_inputs=lin - eps * dm)
    dp = tm.forward_model(shots[0], mp, return_parameters=["simdata"])["simdata"]
    dn = tm.forward_model(shots[0], mm, return_parameters=["simdata"])["simdata"]
    fd = (dp - dn) / (2 * eps)
    assert rel_l2(d1, fd) < 1e-6


def test_dot_product_exact_adjoint_without_pml(gpu):
    """<F m1, d> = <m1, F* d> with prod(deltas) (S carries it, S* does not)."""
    import pysit_amd as pa
    m, shots, solver, model = c1_setup(order=8, tf=0.5, pml=None)
    tm = pa.TemporalModeling(solver)
    rng = np.random.default_rng(4)
    m1 = model.perturbation(data=rng.standard_normal((m.dof(), 1)))
    dW = tm.forward_model(shots[0], model, return_parameters=["dWaveOp"])["dWaveOp"]
    Fm1 = tm.linear_forward_model(shots[0], model, m1, return_parameters=["simdata"], dWaveOp0=dW)["simdata"]
    d = rng.standard_normal(Fm1.shape)
    Fsd = tm.migrate_shot(shots[0], model, d, 1, dWaveOp=dW)
    lhs = solver.dt * float(np.sum(Fm1 * d))
    rhs = float(np.sum(m1.data * Fsd.data)) * np.prod(m.deltas)
    assert abs(lhs - rhs) <= 1e-9 * abs(lhs), (lhs, rhs)


def test_wavefield_returns_and_time_step_api(gpu):
    import pysit_amd as pa
    m, shots, solver, model = c1_setup(order=6, tf=0.08)
    tm = pa.TemporalModeling(solver)
    rv = tm.forward_model(shots[0], model, return_parameters=["wavefield", "simdata"])
    us = rv["wavefield"]
    assert len(us) == solver.nsteps
    k = solver.nsteps - 1
    rec = shots[0].receivers.sample_data_from_array(us[k])
    assert np.allclose(rec.reshape(-1), rv["simdata"][k], rtol=1e-12, atol=1e-300)
    # the per-step API: the reference's own loop shape on solver.time_step
    sd = solver.SolverData()
    rhs = m.pad_array(shots[0].sources.f(0 * solver.dt))
    for step in range(solver.nsteps - 1):
        rhs = m.pad_array(shots[0].sources.f(step * solver.dt))
        solver.time_step(sd, rhs, None)
        sd.advance()
    last = m.unpad_array(sd.k.primary_wavefield)
    assert rel_l2(last, us[solver.nsteps - 1]) < 1e-12


def test_c1_fp32_trace_error(gpu):
    import pysit_amd as pa
    m, shots, solver64, model = c1_setup(order=4)
    solver64.model_parameters = model
    ref = oracle_from_solver(solver64, shots[0]).forward()["simdata"]
    m, shots, solver, model = c1_setup(order=4, precision="single")
    tm = pa.TemporalModeling(solver)
    d = tm.forward_model(shots[0], model, return_parameters=["simdata"])["simdata"]
    err = rel_l2(d, ref)
    print(f"C1 FP32 trace rel L2 vs FP64 oracle: {err:.3e}")
    assert err < 1e-5


def test_lbfgs_two_iterations_decrease_misfit(gpu):
    import pysit_amd as pa
    m, shots, solver, model = c1_setup(order=4, tf=0.8)
    pa.generate_seismic_data(shots, solver, model)
    C0 = model.C.copy()
    C0[:] = 1.0
    m0 = solver.ModelParameters(m, {"C": C0})
    obj = pa.TemporalLeastSquares(solver)
    f0 = obj.evaluate(shots, m0)
    opt = pa.LBFGS(obj, memory_length=5)
    res = opt(shots, m0, 2, status_configuration={"objective_frequency": 1})
    f2 = obj.evaluate(shots, res)
    assert f2 < f0


def test_default_batches_split_at_the_addressable_region(gpu, monkeypatch):
    """With a region that holds one shot, the objective and the data
    generation run one batch per shot by default and give what one batch of
    all shots gives; an explicit batch beyond the region is refused."""
    import pysit_amd as pa
    from pysit_amd import device
    m, _, solver, model = c1_setup(order=4, tf=0.3)
    shots = pa.equispaced_acquisition(m, pa.RickerWavelet(10.0), sources=3, source_depth=0.05,
                                      receivers="max", receiver_depth=0.05)
    pa.generate_seismic_data(shots, solver, model)
    C0 = model.C.copy()
    C0[:] = 1.0
    m0 = solver.ModelParameters(m, {"C": C0})
    obj = pa.TemporalLeastSquares(solver)
    g1, f1 = obj.compute_gradient(shots, m0), obj.evaluate(shots, m0)

    shape = tuple(solver.mesh.shape(include_bc=True, as_grid=True))
    n = device.shot_nodes(shape, 8)
    monkeypatch.setattr(device, "MAX_REGION_BYTES", 56 * n)  # seven FP64 state buffers of one shot
    assert device.shots_per_batch(shape, 8) == 1
    split = copy.deepcopy(shots)
    pa.generate_seismic_data(split, solver, model)
    for a, b in zip(shots, split):
        assert np.array_equal(a.receivers.data, b.receivers.data)
    obj = pa.TemporalLeastSquares(solver)
    g2, f2 = obj.compute_gradient(shots, m0), obj.evaluate(shots, m0)
    assert rel_l2(g2.data.reshape(-1), g1.data.reshape(-1)) < 1e-13
    assert abs(f2 - f1) <= 1e-13 * abs(f1)
    with pytest.raises(ValueError, match="one batch holds at most"):
        pa.TemporalModeling(solver).forward_model_batch(shots, model)
