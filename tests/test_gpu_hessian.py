"""Second-order pieces of TemporalLeastSquares on the GPU (GPU tests).

* The pseudo-Hessian diagonal (aux_info['pseudo_hess_diag'],
  pysit/objective_functions/temporal_least_squares.py:121-139, 169-196): the
  sum over shots and imaged steps of the physical part of dWaveOp^2, times
  imaging_period and dt, against the oracle's dWaveOp history -- with every
  step stored, with checkpoint segments, and with imaging period 3.
* The Gauss-Newton Hessian (apply_hessian 'approximate' and 'levenberg',
  :200-249) is F*F: <m2, H m1> (prod(deltas)-weighted) equals
  dt <F m2, F m1> and <m1, H m2>, with the PML on, for PySIT's and the exact
  dWaveOp; 'levenberg' adds mu m1.
"""
import copy

import numpy as np
import pytest

from tests.problems import c1_setup, oracle_from_solver, rel_l2

pytestmark = pytest.mark.gpu


def _c1_with_data(tf=0.8):
    import pysit_amd as pa
    m, shots, solver, model = c1_setup(order=4, tf=tf)
    shots = copy.deepcopy(shots) + pa.equispaced_acquisition(m, pa.RickerWavelet(10.0), sources=1,
                                                            source_depth=0.3, receivers="max",
                                                            receiver_depth=0.05)
    pa.generate_seismic_data(shots, solver, model)
    X, Z = m.mesh_coords()
    C0 = (1.0 + 0.3 * np.exp(-((Z - 0.45) / 0.02) ** 2)).reshape(-1, 1)
    return m, shots, solver, model, solver.ModelParameters(m, {"C": C0})


@pytest.mark.parametrize("period,budget", [(1, None), (1, 64 << 20), (3, None)])
def test_pseudo_hessian_diagonal_matches_oracle(gpu, period, budget):
    import pysit_amd as pa
    m, shots, solver, model, m0 = _c1_with_data()
    obj = pa.TemporalLeastSquares(solver, imaging_period=period, memory_budget=budget)
    aux = {"pseudo_hess_diag": (True, None), "objective_value": (True, None)}
    obj.compute_gradient(shots, m0, aux_info=aux)
    h = aux["pseudo_hess_diag"][1]
    assert aux["pseudo_hess_diag"][0] and h.shape == (m.dof(), 1)
    if budget:
        assert obj.modeling_tools.last_batch.ckpts, "the budget should force checkpoint segments"
    solver.model_parameters = m0
    ref = np.zeros((m.dof(), 1))
    for shot in shots:
        dw = oracle_from_solver(solver, shot).forward(dwaveop=True, imaging_period=period)["dWaveOp"]
        for a in dw:
            u = m.unpad_array(a.reshape(-1, 1), copy=True)
            ref += u * u
    ref *= period * solver.dt
    err = rel_l2(h, ref)
    print(f"pseudo-Hessian diagonal, period {period}, budget {budget}: rel L2 {err:.2e}")
    assert err < 1e-12


@pytest.mark.parametrize("adjoint", ["pysit", "exact"])
def test_gauss_newton_hessian_is_symmetric_f_star_f(gpu, adjoint):
    import pysit_amd as pa
    m, shots, solver, model = c1_setup(order=4, tf=0.6)
    shot = shots[0]
    obj = pa.TemporalLeastSquares(solver)
    obj.modeling_tools = pa.TemporalModeling(solver, adjoint=adjoint)
    rng = np.random.default_rng(11)
    m1 = model.perturbation(data=rng.standard_normal((m.dof(), 1)))
    m2 = model.perturbation(data=rng.standard_normal((m.dof(), 1)))
    H1 = obj.apply_hessian([shot], model, m1)
    H2 = obj.apply_hessian([shot], model, m2)
    tm = obj.modeling_tools
    dW = tm.forward_model(shot, model, return_parameters=["dWaveOp"])["dWaveOp"]
    F1 = tm.linear_forward_model(shot, model, m1, return_parameters=["simdata"], dWaveOp0=dW)["simdata"]
    F2 = tm.linear_forward_model(shot, model, m2, return_parameters=["simdata"], dWaveOp0=dW)["simdata"]
    w = np.prod(m.deltas)
    a = w * float(np.sum(m2.data * H1.data))
    b = w * float(np.sum(m1.data * H2.data))
    c = solver.dt * float(np.sum(F2 * F1))
    print(f"Gauss-Newton ({adjoint}): <m2,Hm1> {a:.12e}, <m1,Hm2> {b:.12e}, dt<Fm2,Fm1> {c:.12e}")
    assert abs(a - b) <= 1e-9 * abs(c) and abs(a - c) <= 1e-9 * abs(c)
    L = obj.apply_hessian([shot], model, m1, hessian_mode="levenberg", levenberg_mu=0.25)
    assert rel_l2(L.data, H1.data + 0.25 * m1.data) < 1e-13
