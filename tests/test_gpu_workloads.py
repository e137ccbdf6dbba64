"""The benchmark configurations themselves on the GPU (GPU tests).

* 2D layered, 1001 x 301 nodes (1041 x 341 padded), order 8: one shot in
  FP64 against the oracle over 300 steps (traces and gradient to 1e-10),
  and the 16-shot FP32 gradient against the FP64 device path over the full
  3000 steps (traces <= 1e-5, gradient <= 1e-4).
* Uneven shots of the same model (seeded apertures of 300-1001 columns,
  1500-4000 steps): four of them in one FP64 batch against the oracle over
  a tenth of their steps, and eight in FP32 against FP64 over all their
  steps.
* 3D layered, 256^3 nodes (288^3 padded), order 8: one shot in FP64 against
  the oracle over 44 steps, and in FP32 against FP64 over 2000 steps.
* The 2D FWI loop on its full 601 x 201 grid with 32 shots, three bands of
  two LBFGS iterations: misfit and model error decrease in every band.

The FP64 device path matches the oracle (the C restatement of the
reference's FP64 step) to 1e-10, so it stands in for the reference's FP64
run at the full sizes, where the oracle takes minutes.
"""
import numpy as np
import pytest

from tests.problems import oracle_from_geometry, oracle_from_solver, rel_l2, truncated_geometry

pytestmark = pytest.mark.gpu

TRACE_BOUND, GRADIENT_BOUND = 1e-5, 1e-4


def _smoothed(v, dx):
    from scipy.ndimage import gaussian_filter
    v0 = gaussian_filter(v, sigma=300.0 / dx, mode="nearest")
    return v0 * (v.max() / v0.max())


@pytest.mark.timeout(300)
def test_layered_2d_one_shot_fp64_matches_oracle(gpu):
    import pysit_amd as pa
    from pysit_amd.workloads import layered_2d
    mesh, shots, solver, model = layered_2d(n_shots=16, nt=300, precision="double")
    shot = shots[5]
    tm = pa.TemporalModeling(solver, autotune=False)
    rv = tm.forward_model(shot, model, return_parameters=["simdata", "dWaveOp"])
    op = oracle_from_solver(solver, shot)
    ref = op.forward()["simdata"]
    assert np.linalg.norm(ref) > 0
    e_tr = rel_l2(rv["simdata"], ref)
    r = np.random.default_rng(0).standard_normal(ref.shape)
    g = tm.migrate_shot(shot, model, r, 1, dWaveOp=rv["dWaveOp"])
    e_g = rel_l2(g.data.reshape(-1), op.migrate(r))
    print(f"2D layered 1041x341, shot 5, 300 steps, FP64 vs oracle: traces {e_tr:.2e}, gradient {e_g:.2e}")
    assert e_tr < 1e-10 and e_g < 1e-10


def test_layered_2d_16_shots_fp32_within_bounds(gpu):
    """The benchmark's gradient (ShotTraces, checkpoint-capable batch) in
    FP32 against FP64 on the same inputs."""
    import torch
    import pysit_amd as pa
    from pysit_amd.torch_fn import ShotProblem, ShotTraces
    from pysit_amd.workloads import layered_2d
    out = {}
    for prec in ("single", "double"):
        mesh, shots, solver, model = layered_2d(n_shots=16, nt=3000, precision=prec)
        nx, nz = mesh.shape(as_grid=True)
        v = np.asarray(model.C).reshape(nx, nz)
        tm = pa.TemporalModeling(solver, autotune=False)
        d_obs = torch.from_numpy(np.stack(tm.forward_model_batch(shots, model))).cuda()
        tm.last_batch = None
        m = torch.from_numpy(_smoothed(v, mesh.deltas[0]) ** -2.0).cuda().requires_grad_(True)
        prob = ShotProblem(solver, shots, autotune=False)
        tr = ShotTraces.apply(m, prob)
        (0.5 * solver.dt * ((tr - d_obs) ** 2).sum()).backward()
        out[prec] = (tr.detach().cpu().numpy(), m.grad.detach().cpu().numpy())
        del prob, tr, m, d_obs, tm
        torch.cuda.empty_cache()
    e_tr = rel_l2(out["single"][0], out["double"][0])
    e_g = rel_l2(out["single"][1], out["double"][1])
    print(f"2D layered, 16 shots, 3000 steps, FP32 vs FP64: traces {e_tr:.2e}, gradient {e_g:.2e}")
    assert e_tr <= TRACE_BOUND and e_g <= GRADIENT_BOUND


def _uneven(n):
    from pysit_amd.workloads import uneven_2d
    geos, v, deltas, dt, meta = uneven_2d(n_shots=64)
    return geos[:n], v, deltas, dt, meta[:n]


@pytest.mark.timeout(300)
def test_uneven_shots_fp64_batch_matches_oracle(gpu):
    from pysit_amd.device import DeviceBatch
    geos, v, deltas, dt, meta = _uneven(4)
    short = [truncated_geometry(g, g.nt // 10) for g in geos]
    b = DeviceBatch(short, v, deltas, dt, 8, dtype=np.float64)
    traces = b.forward(keep_dw=True)
    rng = np.random.default_rng(2)
    res = [rng.standard_normal((g.nt, g.nrec)) for g in short]
    b.migrate(res)
    for i, g in enumerate(short):
        op = oracle_from_geometry(g, v, deltas, dt, 8)
        ref = op.forward()["simdata"]
        e_tr = rel_l2(traces[i].cpu().numpy(), ref)
        img = b.image(i).cpu().numpy().reshape(-1) * (-dt)
        e_g = rel_l2(img, op.migrate(res[i], fold=False))
        print(f"uneven shot {i} ({meta[i]['w']} columns, {g.nt} steps) vs oracle: traces {e_tr:.2e}, "
              f"image {e_g:.2e}")
        assert e_tr < 1e-10 and e_g < 1e-10


def test_uneven_shots_fp32_within_bounds(gpu):
    import torch
    from pysit_amd.device import DeviceBatch
    geos, v, deltas, dt, meta = _uneven(8)
    out = {}
    for dtype in (np.float32, np.float64):
        b = DeviceBatch(geos, v, deltas, dt, 8, dtype=dtype)
        tr = [t.cpu().numpy().copy() for t in b.forward(keep_dw=True)]
        res = [0.1 * t for t in tr]
        b.migrate(res)
        out[dtype] = (tr, [b.image(i).cpu().numpy().copy() for i in range(len(geos))])
        del b
        torch.cuda.empty_cache()
    worst = (0.0, 0.0)
    for i in range(len(geos)):
        e_tr = rel_l2(out[np.float32][0][i], out[np.float64][0][i])
        e_g = rel_l2(out[np.float32][1][i], out[np.float64][1][i])
        print(f"uneven shot {i} ({meta[i]['w']} columns, {meta[i]['nt']} steps) FP32 vs FP64: traces {e_tr:.2e}, "
              f"image {e_g:.2e}")
        worst = (max(worst[0], e_tr), max(worst[1], e_g))
    assert worst[0] <= TRACE_BOUND and worst[1] <= GRADIENT_BOUND


@pytest.mark.timeout(400)
def test_layered_3d_one_shot_fp64_matches_oracle(gpu):
    # 44 steps: the 6 Hz Ricker rises above its 1e-6 cut-off at step 35, and
    # the receivers next to the source record it from then on
    import pysit_amd as pa
    from pysit_amd.workloads import layered_3d
    mesh, shots, solver, model = layered_3d(n=256, n_shots=1, nt=44, precision="double")
    tm = pa.TemporalModeling(solver, autotune=False)
    rv = tm.forward_model(shots[0], model, return_parameters=["simdata", "dWaveOp"])
    op = oracle_from_solver(solver, shots[0])
    ref = op.forward()["simdata"]
    assert np.linalg.norm(ref) > 0
    e_tr = rel_l2(rv["simdata"], ref)
    r = np.random.default_rng(0).standard_normal(ref.shape)
    g = tm.migrate_shot(shots[0], model, r, 1, dWaveOp=rv["dWaveOp"])
    e_g = rel_l2(g.data.reshape(-1), op.migrate(r))
    print(f"3D layered 288^3, 44 steps, FP64 vs oracle: traces {e_tr:.2e}, gradient {e_g:.2e}")
    assert e_tr < 1e-10 and e_g < 1e-10


def test_layered_3d_one_shot_fp32_within_bounds(gpu):
    import torch
    from pysit_amd.device import DeviceBatch
    from pysit_amd.modeling import TemporalModeling
    from pysit_amd.workloads import layered_3d
    mesh, shots, solver, model = layered_3d(n=256, n_shots=1, nt=2000)
    g = TemporalModeling(solver, autotune=False).geometry(shots[0])
    out = {}
    for dtype in (np.float32, np.float64):
        b = DeviceBatch([g], solver.unpadded_model(), mesh.deltas, solver.dt, 8, dtype=dtype)
        tr = b.forward(keep_dw=True)[0].cpu().numpy().copy()
        b.migrate([0.1 * tr])
        out[dtype] = (tr, b.image(0).cpu().numpy().copy(), b.K)
        del b
        torch.cuda.empty_cache()
    e_tr = rel_l2(out[np.float32][0], out[np.float64][0])
    e_g = rel_l2(out[np.float32][1], out[np.float64][1])
    print(f"3D layered 288^3, 2000 steps, FP32 vs FP64: traces {e_tr:.2e}, image {e_g:.2e} "
          f"(dWaveOp segments of {out[np.float32][2]} and {out[np.float64][2]} slots)")
    assert e_tr <= TRACE_BOUND and e_g <= GRADIENT_BOUND


def test_fwi_loop_full_grid_reduces_misfit_and_model_error(gpu):
    from pysit_amd.workloads import fwi_2d, multiband_fwi
    mesh, solver, true_model, start_model = fwi_2d()
    model, reports = multiband_fwi(mesh, solver, true_model, start_model, iterations=2)
    assert [r["f0_hz"] for r in reports] == [3.0, 5.0, 8.0]
    for r in reports:
        print(f"FWI band {r['f0_hz']} Hz: objective {r['objective'][0]:.4e} -> {r['objective'][-1]:.4e}, "
              f"model error {r['model_error_start']:.4f} -> {r['model_error_final']:.4f}")
        assert r["gradients"] == 2
        assert r["objective"][-1] < r["objective"][0]
        assert r["model_error_final"] <= r["model_error_start"]
    assert reports[-1]["model_error_final"] < reports[0]["model_error_start"]
