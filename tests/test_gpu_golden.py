"""The device path against PySIT's own Python path with the PML on (GPU
tests): tests/golden/pml_golden.npz holds what the reference computes
(tests/golden/make_pml_golden.py) for the 2D horizontal-reflector case and a
3D layered block -- forward_model traces and dWaveOp, migrate_shot of a
closed-form operand, and TemporalLeastSquares' objective and gradient for
data generated in the true model -- and the FP64 device path reproduces them
through this package's API (generate_seismic_data, TemporalModeling,
TemporalLeastSquares) to 1e-10."""
import copy
import os

import numpy as np
import pytest

from tests.problems import golden_operand, pml_golden_case, rel_l2

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pml_golden.npz")


@pytest.mark.parametrize("tag", ["c1", "c3d", "c1o8", "c3do8"])
def test_device_path_matches_reference_python_with_pml(gpu, tag):
    import pysit_amd as pa
    g = np.load(GOLDEN)
    m, shots, solver, C, C0 = pml_golden_case(tag)
    true_model = solver.ModelParameters(m, {"C": C})
    m0 = solver.ModelParameters(m, {"C": C0})
    tm = pa.TemporalModeling(solver, autotune=False)
    rv = tm.forward_model(shots[0], m0, return_parameters=["simdata", "dWaveOp"])
    assert solver.nsteps == int(g[f"{tag}_nsteps"]) and solver.dt == float(g[f"{tag}_dt"])
    every = int(g[f"{tag}_every"])
    e = {"traces": rel_l2(rv["simdata"][::every], g[f"{tag}_traces"])}
    e["dWaveOp"] = max(rel_l2(np.asarray(rv["dWaveOp"][k]).reshape(-1), g[f"{tag}_dwaveop"][i])
                       for i, k in enumerate(g[f"{tag}_dw_steps"]))
    r = golden_operand(*rv["simdata"].shape)
    mig = tm.migrate_shot(shots[0], m0, r, 1, dWaveOp=rv["dWaveOp"])
    e["migrate"] = rel_l2(mig.data.reshape(-1), g[f"{tag}_migrate"])
    obs = copy.deepcopy(shots)
    pa.generate_seismic_data(obs, solver, true_model)
    obj = pa.TemporalLeastSquares(solver)
    e["objective"] = abs(obj.evaluate(obs, m0) - float(g[f"{tag}_objective"])) / float(g[f"{tag}_objective"])
    e["gradient"] = rel_l2(obj.compute_gradient(obs, m0).data.reshape(-1), g[f"{tag}_gradient"])
    print(tag, {k: f"{v:.1e}" for k, v in e.items()})
    assert all(v < 1e-10 for v in e.values()), e


def test_born_hessian_and_lbfgs_match_reference_python_c1(gpu):
    """With the PML on: linear_forward_model traces and the Gauss-Newton
    Hessian (F*F, as apply_hessian composes it) for a closed-form
    perturbation, and three LBFGS iterations (memory 3, backtracking) --
    objective history and final model -- against PySIT's own run."""
    import pysit_amd as pa
    g = np.load(GOLDEN)
    m, shots, solver, C, C0 = pml_golden_case("c1")
    true_model = solver.ModelParameters(m, {"C": C})
    m0 = solver.ModelParameters(m, {"C": C0})
    X, Z = m.mesh_coords()
    m1 = m0.perturbation(data=(0.02 * np.sin(7.0 * X + 1.0) * np.cos(5.0 * Z)).reshape(-1, 1))
    tm = pa.TemporalModeling(solver, autotune=False)
    born = tm.linear_forward_model(shots[0], m0, m1, return_parameters=["simdata"])["simdata"]
    e = {"born": rel_l2(born[::int(g["c1_every"])], g["c1_born"])}
    obs = copy.deepcopy(shots)
    pa.generate_seismic_data(obs, solver, true_model)
    obj = pa.TemporalLeastSquares(solver)
    e["hessian"] = rel_l2(obj.apply_hessian(obs, m0, m1).data.reshape(-1), g["c1_hessian"])
    invalg = pa.LBFGS(obj, memory_length=3)
    res = invalg(obs, m0, 3, line_search="backtrack", status_configuration={"objective_frequency": 1})
    hist = np.asarray([v for _, v in sorted(invalg.objective_history.items())])
    e["lbfgs_objective"] = rel_l2(hist, g["c1_lbfgs_objective"])
    e["lbfgs_model"] = rel_l2(np.asarray(res.C).reshape(-1), g["c1_lbfgs_C"])
    print({k: f"{v:.1e}" for k, v in e.items()})
    assert e["born"] < 1e-10 and e["hessian"] < 1e-10
    assert e["lbfgs_objective"] < 1e-9 and e["lbfgs_model"] < 1e-9, e
