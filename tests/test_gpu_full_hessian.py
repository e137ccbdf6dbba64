"""The full Hessian's second-order pieces on the device (GPU tests):
linear_forward_model's 'dWaveOp1' and 'wavefield1', adjoint_model's
'dWaveOpAdj' and its volume source operand_model * operand_dWaveOpAdj, and
apply_hessian('full') (pysit/objective_functions/temporal_least_squares.py:
224-239, pysit/modeling/temporal_modeling.py:216-226, :338-343, :508-522).

Pinned two ways:

* against PySIT's own run of the same composition with the PML on
  (tests/golden/pml_golden.npz: the two migrations, dWaveOp1 and dWaveOpAdj
  at fixed steps, 2D and 3D) to 1e-10;
* against the mathematics: the derivative of the gradient along m1, by
  central differences of compute_gradient, is res2 - res1 with the PML off
  (both dWaveOp forms coincide there) and, in the exact-adjoint mode, with
  the PML on.  PySIT's 'full' mode adds res1 instead (its residual enters
  with the opposite sign of the gradient's): the test records that the sum
  is off the derivative by the size of res1.
"""
import copy
import os

import numpy as np
import pytest

from tests.problems import c1_setup, golden_direction, pml_golden_case, rel_l2

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pml_golden.npz")


@pytest.mark.parametrize("tag", ["c1", "c3d"])
def test_full_hessian_terms_match_reference_python(gpu, tag):
    import pysit_amd as pa
    g = np.load(GOLDEN)
    m, shots, solver, C, C0 = pml_golden_case(tag)
    true_model = solver.ModelParameters(m, {"C": C})
    m0 = solver.ModelParameters(m, {"C": C0})
    m1 = m0.perturbation(data=golden_direction(m))
    obs = copy.deepcopy(shots)
    pa.generate_seismic_data(obs, solver, true_model)
    obj = pa.TemporalLeastSquares(solver)
    tm = obj.modeling_tools
    # the pieces through the modeling API, as the reference composes them
    dw0 = tm.forward_model(obs[0], m0, 1, return_parameters=["simdata", "dWaveOp"])
    r0 = obs[0].receivers.interpolate_data(solver.ts()) - dw0["simdata"]
    lin = tm.linear_forward_model(obs[0], m0, m1, return_parameters=["simdata", "dWaveOp1", "wavefield1"],
                                  dWaveOp0=dw0["dWaveOp"])
    adj1 = []
    res1 = tm.migrate_shot(obs[0], m0, r0, 1, dWaveOp=lin["dWaveOp1"], dWaveOpAdj=adj1)
    e = {"born": rel_l2(lin["simdata"][::int(g[f"{tag}_every"])], g[f"{tag}_born"])}
    steps = g[f"{tag}_dw_steps"]
    e["dWaveOp1"] = max(rel_l2(np.asarray(lin["dWaveOp1"][k]).reshape(-1), g[f"{tag}_dwaveop1"][i])
                        for i, k in enumerate(steps))
    e["dWaveOpAdj"] = max(rel_l2(np.asarray(adj1[k]).reshape(-1), g[f"{tag}_dwaveopadj"][i])
                          for i, k in enumerate(steps))
    res2 = tm.migrate_shot(obs[0], m0, lin["simdata"], 1, operand_dWaveOpAdj=adj1, operand_model=m1,
                           dWaveOp=dw0["dWaveOp"])
    e["res1"] = rel_l2(res1.data.reshape(-1), g[f"{tag}_full_res1"])
    e["res2"] = rel_l2(res2.data.reshape(-1), g[f"{tag}_full_res2"])
    full = obj.apply_hessian(obs, m0, m1, hessian_mode="full")
    e["full"] = rel_l2(full.data.reshape(-1), g[f"{tag}_full_res1"] + g[f"{tag}_full_res2"])
    # wavefield1 is the Born field the traces sample: its last step's
    # receiver samples are the last Born trace row
    k = solver.nsteps - 1
    rec = obs[0].receivers.sample_data_from_array(lin["wavefield1"][k])
    e["wavefield1"] = rel_l2(np.asarray(rec).reshape(-1), lin["simdata"][k])
    print(tag, {key: f"{v:.1e}" for key, v in e.items()})
    assert all(v < 1e-10 for v in e.values()), e


def _fd_case(adjoint, pml):
    """C1 at T = 0.6 with a bump perturbation away from the fast reflector
    (the largest wave speed, hence dt and the step count, stays put under
    +-eps m1), data generated in the true model, start model with the
    reflector at 1.3.  The bump is below 1e-11 on the boundary of the
    physical grid: the Born source pads m1 by edge into the PML
    (temporal_modeling.py:432), the adjoint's volume source pads it with
    zeros (:310), and the two agree only where m1 vanishes at the boundary."""
    import pysit_amd as pa
    m, shots, solver, model = c1_setup(order=4, tf=0.6, pml=pml)
    pa.generate_seismic_data(shots, solver, model)
    X, Z = m.mesh_coords()
    C0 = (1.0 + 0.3 * np.exp(-((Z - 0.45) / 0.02) ** 2)).reshape(-1, 1)
    m0 = solver.ModelParameters(m, {"C": C0})
    dm = (0.05 * np.exp(-((X - 0.5) ** 2 + (Z - 0.3) ** 2) / 0.0035) * (Z < 0.36)).reshape(-1, 1)
    m1 = m0.perturbation(data=dm)
    obj = pa.TemporalLeastSquares(solver)
    obj.modeling_tools = pa.TemporalModeling(solver, adjoint=adjoint)
    return m, shots, solver, m0, m1, obj


def _fd_gradient_derivative(solver, shots, m0, m1, obj, eps=1e-4):
    m = solver.mesh
    lin = m0.linearize()
    solver.model_parameters = m0
    nsteps = solver.nsteps
    g = []
    for sign in (1, -1):
        mp = solver.ModelParameters(m, linear_inputs=lin + sign * eps * m1.data)
        solver.model_parameters = mp
        assert solver.nsteps == nsteps  # the time axis stays put
        g.append(obj.compute_gradient(shots, mp).data.reshape(-1))
    return (g[0] - g[1]) / (2 * eps)


@pytest.mark.parametrize("adjoint,pml", [("pysit", None), ("exact", 0.1)])
def test_gradient_derivative_is_res2_minus_res1(gpu, adjoint, pml):
    m, shots, solver, m0, m1, obj = _fd_case(adjoint, pml)
    fd = _fd_gradient_derivative(solver, shots, m0, m1, obj)
    res1, res2 = obj.full_hessian_terms(shots[0], m0, m1)
    r1, r2 = res1.data.reshape(-1), res2.data.reshape(-1)
    e_exact = rel_l2(r2 - r1, fd)
    e_pysit = rel_l2(r2 + r1, fd)
    e_gn = rel_l2(obj.apply_hessian(shots, m0, m1).data.reshape(-1), fd)
    size1 = np.linalg.norm(r1) / np.linalg.norm(fd)
    print(f"{adjoint} pml={pml}: |res1|/|dg| {size1:.2e}; res2-res1 {e_exact:.2e}, "
          f"res1+res2 {e_pysit:.2e}, Gauss-Newton {e_gn:.2e}", flush=True)
    assert e_exact < 1e-7
    # the reference's sum is off by about 2 |res1|
    assert e_pysit > 0.5 * size1


def test_adjoint_without_operand_model_takes_no_volume_source(gpu):
    """The reference adds the volume source only when operand_model is given
    too (temporal_modeling.py:221-222)."""
    import pysit_amd as pa
    m, shots, solver, model = c1_setup(order=4, tf=0.3)
    tm = pa.TemporalModeling(solver)
    dW = tm.forward_model(shots[0], model, 1, return_parameters=["dWaveOp"])["dWaveOp"]
    r = np.random.default_rng(5).standard_normal((solver.nsteps, shots[0].receivers.receiver_count))
    adj = []
    a = tm.migrate_shot(shots[0], model, r, 1, dWaveOp=dW, dWaveOpAdj=adj)
    b = tm.migrate_shot(shots[0], model, r, 1, dWaveOp=dW, operand_dWaveOpAdj=adj)
    c = tm.migrate_shot(shots[0], model, r, 1, dWaveOp=dW)
    assert np.array_equal(a.data, c.data) and np.array_equal(b.data, c.data)
    # host arrays give what the device handles give
    host = [np.asarray(h) for h in adj]
    m1 = model.perturbation(data=np.random.default_rng(6).standard_normal((m.dof(), 1)))
    d = tm.migrate_shot(shots[0], model, r, 1, dWaveOp=dW, operand_dWaveOpAdj=adj, operand_model=m1)
    e = tm.migrate_shot(shots[0], model, r, 1, dWaveOp=dW, operand_dWaveOpAdj=host, operand_model=m1)
    assert np.array_equal(d.data, e.data)
    assert rel_l2(d.data, c.data) > 1e-6


@pytest.mark.parametrize("tag", ["c1", "c3d"])
def test_full_hessian_fp32_follows_fp64(gpu, tag):
    """The FP32 second-order runs (Born storing dWaveOp1, the adjoint storing
    dWaveOpAdj and taking the volume source) against the FP64 ones."""
    import pysit_amd as pa
    out = {}
    for prec in ("double", "single"):
        m, shots, solver, C, C0 = pml_golden_case(tag, precision=prec)
        m0 = solver.ModelParameters(m, {"C": C0})
        obs = copy.deepcopy(shots)
        pa.generate_seismic_data(obs, solver, solver.ModelParameters(m, {"C": C}))
        obj = pa.TemporalLeastSquares(solver)
        r1, r2 = obj.full_hessian_terms(obs[0], m0, m0.perturbation(data=golden_direction(m)))
        out[prec] = (r1.data.reshape(-1), r2.data.reshape(-1))
    e1 = rel_l2(out["single"][0], out["double"][0])
    e2 = rel_l2(out["single"][1], out["double"][1])
    print(tag, f"FP32 vs FP64: res1 {e1:.1e}, res2 {e2:.1e}")
    assert e1 < 1e-4 and e2 < 1e-5
