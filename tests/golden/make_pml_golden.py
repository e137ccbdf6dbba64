"""Write golden vectors of PySIT's own Python path with the PML on.

The steps are pinned bit for bit by step2d/step3d.npz; what only PySIT's
Python can pin is their composition with the PML on: forward_model
(pysit/modeling/temporal_modeling.py:60-169), migrate_shot (:171-214) with
adjoint_model (:228-387) and the PML fold (solvers/model_parameter.py:
537-593), generate_seismic_data, and TemporalLeastSquares
(objective_functions/temporal_least_squares.py:27-198).  This script imports
the reference from /root/reference in the container that holds it (never on
the GPU host) and writes ``tests/golden/pml_golden.npz``.

Run after ``make -C oracle refpy``, which compiles the reference's SWIG
wrapper (the shipped solvers_wrap.cxx) in place into ``oracle/_ref``:

    python tests/golden/make_pml_golden.py

Shims, none of them on the computed path: numpy 2 lacks the ``np.int`` (and
float, complex, bool) aliases the reference's point operators use; pyamg is
absent, and ``pyamg.gallery.stencil_grid`` / ``pyamg.krylov`` are only
reached by the frequency-domain and Gauss-Newton code, so stubs that raise
stand in for them; the extension is loaded under its package name from
oracle/_ref.

Cases (closed-form inputs, no random generator in the fixture's recipe):

* ``c1_*`` -- the 2D horizontal reflector (x [0, 0.9], z [0, 0.7], 91 x 71,
  PML(0.1, 100) on all sides, order 4, T = 1.5, one 10 Hz Ricker at depth
  0.05, 91 receivers at depth 0.05, 'cpp' kernel, FP64): traces of the start
  model (the reflector at 1.3 instead of 1.5) every 4th step, its dWaveOp at
  40 % and 90 % of the steps, migrate_shot of the closed-form operand
  r[k, j] = sin(0.37 k + 1.3 j) exp(-k / 800), and TemporalLeastSquares'
  objective and gradient at the start model for data generated in the true
  model; then, at the start model, linear_forward_model (Born) traces,
  apply_hessian ('approximate') and the two migrations of apply_hessian
  ('full') with the dWaveOp of the Born and of the adjoint field at the same
  steps, for the closed-form perturbation m1 = 0.02 sin(7x + 1) cos(5z)
  (0.02 sin(7x + 1) cos(5z) cos(3y) in 3D), and the objective history and
  final model of three LBFGS iterations (memory 3, backtracking line search,
  optimization.py:312-501, lbfgs.py:39-124).
* ``c1o8_*``, ``c3do8_*`` -- the same two cases at spatial order 8: traces,
  dWaveOp, migrate_shot, objective and gradient.
* ``c3d_*`` -- a 3D layered block (x, y, z [0, 0.3], 16^3 nodes, PML(0.08,
  60) on all sides, order 4, T = 0.25, one 12 Hz Ricker, a 4 x 4 receiver
  patch), the same quantities but LBFGS: traces every 2nd step, dWaveOp at
  70 % of the steps, migrate_shot of the same closed-form operand, objective
  and gradient, Born traces and both Hessians.
"""
import glob
import importlib.util
import os
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"


def import_reference():
    for name, val in (("int", int), ("float", float), ("complex", complex), ("bool", bool)):
        if not hasattr(np, name):
            setattr(np, name, val)

    def absent(*a, **k):
        raise RuntimeError("pyamg is absent; the time-domain path does not use it")

    pa = types.ModuleType("pyamg")
    pa.__path__ = []
    pg = types.ModuleType("pyamg.gallery")
    pk = types.ModuleType("pyamg.krylov")
    pg.stencil_grid = absent
    pk.cg = pk.gmres = absent
    pa.gallery, pa.krylov = pg, pk
    sys.modules.update({"pyamg": pa, "pyamg.gallery": pg, "pyamg.krylov": pk})
    full = "pysit.solvers.constant_density_acoustic.time.scalar._constant_density_acoustic_time_scalar_cpp"
    so = glob.glob(os.path.join(ROOT, "oracle", "_ref", "_constant_density_acoustic_time_scalar_cpp*.so"))
    if not so:
        raise SystemExit("the reference extension is not built (make -C oracle refpy)")
    spec = importlib.util.spec_from_file_location(full, so[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    sys.modules[full] = mod
    sys.path.insert(0, REF)
    import pysit
    return pysit


def operand(nt, nrec):
    k = np.arange(nt)[:, None]
    j = np.arange(nrec)[None, :]
    return np.sin(0.37 * k + 1.3 * j) * np.exp(-k / 800.0)


def run_case(ps, m, shots, solver, C, C0, every, dw_steps, second_order=False, lbfgs=False):
    import copy
    out = {}
    true_model = solver.ModelParameters(m, {"C": C})
    m0 = solver.ModelParameters(m, {"C": C0})
    tools = ps.TemporalModeling(solver)
    rv = tools.forward_model(shots[0], m0, return_parameters=["simdata", "dWaveOp"])
    out["traces"] = np.asarray(rv["simdata"])[::every].copy()
    dw_steps = [int(f * solver.nsteps) for f in dw_steps]
    out["dwaveop"] = np.stack([np.asarray(rv["dWaveOp"][k]).reshape(-1) for k in dw_steps])
    out["dw_steps"] = np.asarray(dw_steps)
    r = operand(*np.asarray(rv["simdata"]).shape)
    g = tools.migrate_shot(shots[0], m0, r, 1, dWaveOp=rv["dWaveOp"])
    out["migrate"] = np.asarray(g.data).reshape(-1).copy()
    obs = copy.deepcopy(shots)
    ps.generate_seismic_data(obs, solver, true_model)
    obj = ps.TemporalLeastSquares(solver)
    out["objective"] = np.asarray(obj.evaluate(obs, m0))
    out["gradient"] = np.asarray(obj.compute_gradient(obs, m0).data).reshape(-1).copy()
    out["nsteps"] = np.asarray(solver.nsteps)
    out["dt"] = np.asarray(solver.dt)
    out["every"] = np.asarray(every)
    if second_order:
        # Born (linear_forward_model) and the Gauss-Newton Hessian applied to
        # a closed-form perturbation, then three LBFGS iterations (memory 3,
        # backtracking) from the start model
        m1 = m0.perturbation(data=born_direction(m))
        lin = tools.linear_forward_model(shots[0], m0, m1, return_parameters=["simdata"])
        out["born"] = np.asarray(lin["simdata"])[::every].copy()
        # apply_hessian('approximate') as the reference composes it
        # (temporal_least_squares.py:214-222): its own call of migrate_shot
        # omits the required imaging_period argument (temporal_modeling.py:171)
        # and raises TypeError, so the composition is spelled out with
        # imaging_period 1
        dw0 = tools.forward_model(obs[0], m0, return_parameters=["dWaveOp"])["dWaveOp"]
        d1 = tools.linear_forward_model(obs[0], m0, m1, return_parameters=["simdata"], dWaveOp0=dw0)["simdata"]
        out["hessian"] = np.asarray(tools.migrate_shot(obs[0], m0, d1, 1, dWaveOp=dw0).data).reshape(-1).copy()
        # apply_hessian('full') (temporal_least_squares.py:224-239), spelled
        # out for the same reason: the Born field's dWaveOp1, the adjoint of
        # the residual imaged against it keeping its own dWaveOp (dWaveOpAdj),
        # and the adjoint of the Born traces with the volume source
        # m1 * dWaveOpAdj_k imaged against dWaveOp0
        dw0 = []
        r0 = obj._residual(obs[0], m0, dWaveOp=dw0)
        lin = tools.linear_forward_model(obs[0], m0, m1, return_parameters=["simdata", "dWaveOp1"], dWaveOp0=dw0)
        adj1 = []
        res1 = tools.migrate_shot(obs[0], m0, r0, 1, dWaveOp=lin["dWaveOp1"], dWaveOpAdj=adj1)
        res2 = tools.migrate_shot(obs[0], m0, lin["simdata"], 1, operand_dWaveOpAdj=adj1, operand_model=m1,
                                  dWaveOp=dw0)
        out["full_res1"] = np.asarray(res1.data).reshape(-1).copy()
        out["full_res2"] = np.asarray(res2.data).reshape(-1).copy()
        out["dwaveop1"] = np.stack([np.asarray(lin["dWaveOp1"][k]).reshape(-1) for k in dw_steps])
        out["dwaveopadj"] = np.stack([np.asarray(adj1[k]).reshape(-1) for k in dw_steps])
    if lbfgs:
        invalg = ps.LBFGS(obj, memory_length=3)
        res = invalg(obs, m0, 3, line_search="backtrack", status_configuration={"objective_frequency": 1})
        out["lbfgs_objective"] = np.asarray([v for _, v in sorted(invalg.objective_history.items())])
        out["lbfgs_C"] = np.asarray(res.C).reshape(-1).copy()
    return out


def born_direction(m):
    """A smooth closed-form model perturbation over the whole physical grid."""
    if m.dim == 3:
        X, Y, Z = m.mesh_coords()
        return (0.02 * np.sin(7.0 * X + 1.0) * np.cos(5.0 * Z) * np.cos(3.0 * Y)).reshape(-1, 1)
    X, Z = m.mesh_coords()
    return (0.02 * np.sin(7.0 * X + 1.0) * np.cos(5.0 * Z)).reshape(-1, 1)


def c1(ps, order=4):
    p = ps.PML(0.1, 100)
    d = ps.RectangularDomain((0.0, 0.9, p, p), (0.0, 0.7, p, p))
    m = ps.CartesianMesh(d, 91, 71)
    X, Z = m.mesh_coords()
    C = (1.0 + 0.5 * np.exp(-((Z - 0.45) / 0.02) ** 2)).reshape(-1, 1)
    C0 = (1.0 + 0.3 * np.exp(-((Z - 0.45) / 0.02) ** 2)).reshape(-1, 1)
    shots = ps.equispaced_acquisition(m, ps.RickerWavelet(10.0), sources=1, source_depth=0.05,
                                      receivers="max", receiver_depth=0.05)
    solver = ps.ConstantDensityAcousticWave(m, spatial_accuracy_order=order, trange=(0.0, 1.5),
                                            kernel_implementation="cpp")
    if order != 4:
        return run_case(ps, m, shots, solver, C, C0, 4, [0.4, 0.9])
    return run_case(ps, m, shots, solver, C, C0, 4, [0.4, 0.9], second_order=True, lbfgs=True)


def block3d(ps, order=4):
    p = ps.PML(0.08, 60)
    d = ps.RectangularDomain((0.0, 0.3, p, p), (0.0, 0.3, p, p), (0.0, 0.3, p, p))
    m = ps.CartesianMesh(d, 16, 16, 16)
    X, Y, Z = m.mesh_coords()
    C = (1.0 + 0.4 * (Z > 0.17) + 0.15 * np.exp(-((X - 0.15) ** 2 + (Y - 0.15) ** 2 + (Z - 0.1) ** 2) / 0.004))
    C0 = 1.0 + 0.4 * (Z > 0.17)
    C, C0 = C.reshape(-1, 1), C0.reshape(-1, 1)
    rec = np.linspace(0.05, 0.25, 4)
    src = ps.PointSource(m, (0.15, 0.15, 0.04), ps.RickerWavelet(12.0))
    recs = ps.ReceiverSet(m, [ps.PointReceiver(m, (a, b, 0.04)) for a in rec for b in rec])
    shots = [ps.Shot(src, recs)]
    solver = ps.ConstantDensityAcousticWave(m, spatial_accuracy_order=order, trange=(0.0, 0.25),
                                            kernel_implementation="cpp")
    return run_case(ps, m, shots, solver, C, C0, 2, [0.7], second_order=order == 4)


def main():
    ps = import_reference()
    out = {}
    cases = (("c1", c1), ("c3d", block3d), ("c1o8", lambda ps: c1(ps, 8)), ("c3do8", lambda ps: block3d(ps, 8)))
    for tag, fn in cases:
        res = fn(ps)
        print(tag, {k: np.shape(v) for k, v in res.items()}, flush=True)
        out.update({f"{tag}_{k}": v for k, v in res.items()})
    path = os.path.join(HERE, "pml_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
