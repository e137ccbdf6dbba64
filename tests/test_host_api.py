"""Host-side pieces of the drop-in API (CPU tests): closed forms for the
geometry, PML, point operators, wavelet and CFL step, the factory contract,
model arithmetic, and the C ABI library's exports."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import pysit_amd as pa
from tests.problems import c1_setup

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_pml_profile_closed_form():
    # pysit/core/mesh.py:749-782 with domain.py:236-254: n = ceil(L/d);
    # left[j] = A((n-j)/n)^2, right[j] = A((j+1)/n)^2
    p = pa.PML(0.1, 100.0)
    d = pa.RectangularDomain((0.0, 0.9, p, p), (0.0, 0.7, p, p))
    m = pa.CartesianMesh(d, 91, 71)
    n = m.x.lbc.n
    assert n == 10 and m.z.rbc.n == 10
    j = np.arange(n)
    assert np.allclose(m.x.lbc.sigma, 100.0 * ((n - j) / n) ** 2, rtol=1e-15)
    assert np.allclose(m.x.rbc.sigma, 100.0 * ((j + 1) / n) ** 2, rtol=1e-15)
    assert m.shape(include_bc=True, as_grid=True) == [111, 91]


def test_cfl_dt_and_nsteps_c1():
    m, shots, solver, model = c1_setup(order=4)
    solver.model_parameters = model
    assert abs(solver.dt - 0.01 / (6 * 1.5)) < 1e-15
    assert solver.nsteps == 1350
    assert np.allclose(solver.ts()[:3], [0.0, solver.dt, 2 * solver.dt])


def test_edge_padding_of_the_model():
    m, shots, solver, model = c1_setup(order=4)
    solver.model_parameters = model
    Cp = solver.model_parameters.C.reshape(111, 91)
    C = model.C.reshape(91, 71)
    assert np.array_equal(Cp[10:101, 10:81], C)
    assert np.array_equal(Cp[0, 10:81], C[0])
    assert np.array_equal(Cp[:10, 0], np.full(10, C[0, 0]))


def test_gaussian_point_operator():
    m, shots, solver, model = c1_setup(order=4)
    src = shots[0].sources
    idx, w = src.entries()
    assert idx.size == 49
    # the weights integrate to one over the mesh, up to the 3-sigma truncation
    assert abs(w.sum() * np.prod(m.deltas) - 1.0) < 2e-3
    # closed form at the centre node: 1/(2 pi sigma^2) * exp(-|x-p|^2 / 2 sigma^2)
    X, Z = m.mesh_coords()
    px, pz = src.position
    flat = np.ravel_multi_index((int(round(px / 0.01)), int(round(pz / 0.01))), (91, 71))
    k = list(idx).index(flat)
    want = 1.0 / (2 * np.pi * 0.01 ** 2) * np.exp(-((X.reshape(-1)[flat] - px) ** 2 + (Z.reshape(-1)[flat] - pz) ** 2)
                                                  / (2 * 0.01 ** 2))
    assert abs(w[k] - want) <= 1e-12 * want
    # S = w * prod(deltas), S* = w
    rec = shots[0].receivers
    u = np.random.default_rng(0).standard_normal((m.dof(), 1))
    d = rec.sample_data_from_array(u)
    r0 = rec.receiver_list[0]
    i0, w0 = r0.entries()
    assert abs(d[0, 0] - np.sum(w0 * u[i0, 0]) * np.prod(m.deltas)) < 1e-12


def test_delta_point_operator():
    m, shots, solver, model = c1_setup(order=4)
    s = pa.PointSource(m, (0.3, 0.2), pa.RickerWavelet(5.0), approximation="delta")
    idx, w = s.entries()
    assert idx.size == 1 and abs(w[0] - 1.0 / np.prod(m.deltas)) < 1e-9


def test_ricker_closed_form():
    w = pa.RickerWavelet(10.0)
    sigma = 1.0 / (math.pi * 10.0 * math.sqrt(2.0))
    t = np.linspace(0.0, 0.3, 301)
    x = (t - 6 * sigma) / (math.sqrt(2.0) * sigma)
    want = (1 - 2 * x ** 2) * np.exp(-x ** 2)
    want[np.abs(want) < 1e-6] = 0.0
    got = np.array([w(tt) for tt in t])
    assert np.allclose(got, want, rtol=1e-12, atol=1e-15)
    assert abs(w(6 * sigma) - 1.0) < 1e-14


def test_factory_dispatch():
    m, shots, solver, model = c1_setup(order=4)
    s = pa.ConstantDensityAcousticWave(m, spatial_accuracy_order=8, kernel_implementation="hip")
    assert type(s).__name__ == "ConstantDensityAcousticTimeScalar_2D_hip"
    assert s.supports["spatial_dimension"] == 2 and s.supports["temporal_integrator"] == "leap-frog"
    s = pa.ConstantDensityAcousticWave(m)  # first partial match
    assert type(s).__name__ == "ConstantDensityAcousticTimeScalar_2D_hip"
    s = pa.ConstantDensityAcousticWave(m, spatial_accuracy_order=6, some_unknown_kwarg=3)
    assert s.spatial_accuracy_order == 6
    for bad in ({"spatial_accuracy_order": 10}, {"kernel_implementation": "cpp"},
                {"equation_dynamics": "frequency"}):
        with pytest.raises(pa.NoMatchError):
            pa.ConstantDensityAcousticWave(m, **bad)
    p = pa.PML(0.1, 10.0)
    d3 = pa.RectangularDomain((0.0, 0.1, p, p), (0.0, 0.1, p, p), (0.0, 0.1, p, p))
    m3 = pa.CartesianMesh(d3, 11, 11, 11)
    s3 = pa.ConstantDensityAcousticWave(m3, spatial_accuracy_order=4)
    assert type(s3).__name__ == "ConstantDensityAcousticTimeScalar_3D_hip"
    assert s3.WavefieldVector.aux_names == ["psi", "Phix", "Phiy", "Phiz"]


def test_model_arithmetic_is_linear_in_slowness_squared():
    m, shots, solver, model = c1_setup(order=4)
    p = model.perturbation(data=np.full((m.dof(), 1), 0.1))
    m2 = model + p
    assert np.allclose(m2.C, (model.C ** -2 + 0.1) ** -0.5)
    diff = m2 - model
    assert isinstance(diff, type(p)) and np.allclose(diff.data, 0.1)
    assert abs(p.inner_product(p) - 0.01 * m.dof() * np.prod(m.deltas)) < 1e-12
    assert model.validate()
    bad = solver.ModelParameters(m, {"C": -model.C})
    assert not bad.validate()


def test_host_fold_is_transpose_of_edge_padding():
    m, shots, solver, model = c1_setup(order=4)
    rng = np.random.default_rng(0)
    x = model.perturbation(data=rng.standard_normal((m.dof(), 1)))
    y = rng.standard_normal((m.dof(include_bc=True), 1))
    lhs = float(np.sum(x.with_padding(padding_mode="edge").data * y))
    P = x.perturbation if hasattr(x, "perturbation") else None
    yp = type(x)(m, padded=True, inputs=y)
    rhs = float(np.sum(x.data * yp.add_padding().data))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_receiver_interpolation_zero_fill():
    m, shots, solver, model = c1_setup(order=4)
    rec = shots[0].receivers
    ts = np.arange(11) * 0.1
    rec.reset_time_series(ts)
    rec.data = np.outer(2 * ts + 1, np.ones(rec.receiver_count))
    tn = np.array([-0.05, 0.0, 0.333, 1.0, 1.2])
    d = rec.interpolate_data(tn)
    want = np.where((tn >= 0) & (tn <= 1.0), 2 * tn + 1, 0.0)
    assert np.allclose(d[:, 3], want, atol=1e-14)


def test_c_abi_exports_every_declared_symbol():
    """libpysit_amd.so exports each entry point include/pysit_amd.h declares
    (loaded, no device calls)."""
    lib_path = os.path.join(ROOT, "pysit_amd", "libpysit_amd.so")
    if not os.path.exists(lib_path):
        pytest.skip("library not built (python -m pysit_amd._build)")
    header = open(os.path.join(ROOT, "include", "pysit_amd.h")).read()
    declared = set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pysit_amd_\w+)\s*\(", header, re.M))
    assert declared >= {"pysit_amd_batch_run", "pysit_amd_cda_step2d_f32", "pysit_amd_fold"}
    L = ctypes.CDLL(lib_path)
    for name in declared:
        assert hasattr(L, name), name
    from pysit_amd import _native
    assert declared == set(_native.exported_symbols())
    L.pysit_amd_abi_check.restype = ctypes.c_int
    assert L.pysit_amd_abi_check(ctypes.sizeof(_native.Shot), ctypes.sizeof(_native.Batch)) == 0


def test_solver_path_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    m, shots, solver, model = c1_setup(order=4, tf=0.05)
    tm = pa.TemporalModeling(solver)
    with pytest.raises(Exception) as e:
        tm.forward_model(shots[0], model, return_parameters=["simdata"])
    assert "GPU" in str(e.value) or "ROCm" in str(e.value)


def test_plan_entry_points_check_arguments_without_a_device():
    """pysit_amd_plan_create rejects bad sizes before touching a device; the
    other plan entry points reject a null plan."""
    from pysit_amd import _native as N
    if not os.path.exists(N.LIB_PATH):
        pytest.skip("library not built (python -m pysit_amd._build)")
    L = N.lib()
    h = ctypes.c_void_p()
    for gs, mg, what in ((0, 4, "graph_steps"), (5000, 4, "graph_steps"), (32, 0, "max_graphs")):
        assert L.pysit_amd_plan_create(gs, mg, ctypes.byref(h)) != 0
        assert what in L.pysit_amd_last_error().decode()
        assert h.value is None
    assert L.pysit_amd_plan_run(None, None, 0, 0, 1, 1, 0, None, None, None) != 0
    assert "null plan" in L.pysit_amd_last_error().decode()
    stats = (ctypes.c_longlong * 4)()
    assert L.pysit_amd_plan_stats(None, ctypes.cast(stats, ctypes.c_void_p)) != 0
    assert L.pysit_amd_plan_destroy(None) == 0


def _c1_shots_with_data(n=3, seed=0):
    import pysit_amd as pa
    m, shots, solver, model = c1_setup(order=4)
    shots = pa.equispaced_acquisition(m, pa.RickerWavelet(10.0), sources=n, source_depth=0.05, receivers="max",
                                      receiver_depth=0.05)
    rng = np.random.default_rng(seed)
    ts = np.arange(50) * 0.001
    for sh in shots:
        sh.receivers.reset_time_series(ts)
        sh.receivers.data = rng.standard_normal((50, sh.receivers.receiver_count))
    return m, shots


def test_supershot_encodes_sources_and_data():
    """SourceEncodedSupershot (shot.py:149-344): the encoded data are
    sum_i c_i d_i, the encoded sources carry the codes as intensities, the
    injection has one channel per member source scaled by its code."""
    import pysit_amd as pa
    m, shots = _c1_shots_with_data(3)
    sup = pa.SourceEncodedSupershot(shots, rng=np.random.default_rng(5))
    c = sup.codes
    assert c.shape == (3,)
    want = sum(ci * sh.receivers.data for ci, sh in zip(c, shots))
    assert np.array_equal(sup.receivers.data, want)
    assert np.array_equal(sup.receivers.ts, shots[0].receivers.ts)
    assert [s.intensity for s in sup.sources.source_list] == list(c)
    assert all(sh.sources.intensity == 1.0 for sh in shots)  # members untouched
    (pos, chan, w), amp = sup.sources.injection(0.001, 40)
    (_, _, _), amp0 = shots[1].sources.injection(0.001, 40)
    assert amp.shape == (40, 3) and np.allclose(amp[:, 1], c[1] * amp0[:, 0], rtol=0, atol=1e-15)
    sup.encode()
    assert not np.array_equal(sup.codes, c)
    k = pa.SourceEncodedSupershot(shots, weight_type="krebs", rng=np.random.default_rng(1))
    assert set(np.unique(k.codes)) <= {-1.0, 1.0}


def test_supershot_rejects_mixed_acquisition():
    import pysit_amd as pa
    m, shots = _c1_shots_with_data(2)
    other = pa.equispaced_acquisition(m, pa.RickerWavelet(10.0), sources=1, source_depth=0.05, receivers=45,
                                      receiver_depth=0.05)[0]
    other.receivers.reset_time_series(shots[0].receivers.ts)
    other.receivers.data = np.zeros((50, 45))
    with pytest.raises(Exception, match="same for each shot"):
        pa.SourceEncodedSupershot([shots[0], other])
    with pytest.raises(Exception, match="weight type"):
        pa.SourceEncodedSupershot(shots, weight_type="uniform")


@pytest.mark.parametrize("method", ["savemat", "pickle", "npy"])
def test_shot_data_files_round_trip(method, tmp_path, monkeypatch):
    """Shot data on disk (data_modeling.py:62-131 save, :180-330 load): each
    shot's traces and time axis come back bit for bit into fresh shots'
    receivers, under ./shots_time/shot_<i>, ts_<i>; a save removes the shot
    files of an earlier one and leaves other files alone."""
    import pysit_amd as pa
    from pysit_amd.modeling.data_modeling import _EXT
    monkeypatch.chdir(tmp_path)
    m, shots = _c1_shots_with_data(3, seed=1)
    solver = pa.ConstantDensityAcousticWave(m, spatial_accuracy_order=4)
    (tmp_path / "shots_time").mkdir()
    (tmp_path / "shots_time" / "notes.txt").write_text("x")
    for old in ("shot_7.mat", "ts_7.mat", "shot_4.save", "ts_9.npy"):
        (tmp_path / "shots_time" / old).write_text("stale")
    pa.save_shot_data(shots, method)
    names = sorted(os.listdir(tmp_path / "shots_time"))
    ext = _EXT[method]
    assert names == sorted([f"shot_{i}{ext}" for i in (1, 2, 3)] + [f"ts_{i}{ext}" for i in (1, 2, 3)] +
                           ["notes.txt"])
    _, fresh = _c1_shots_with_data(3, seed=2)
    pa.generate_seismic_data_from_file(fresh, solver, save_method=method)
    for a, b in zip(shots, fresh):
        assert np.array_equal(np.asarray(b.receivers.ts), np.asarray(a.receivers.ts))
        assert np.array_equal(np.asarray(b.receivers.data), np.asarray(a.receivers.data))
        d = b.receivers.interpolate_data(a.receivers.ts) - a.receivers.data
        assert np.max(np.abs(d)) <= 1e-13 * np.max(np.abs(a.receivers.data))


def test_shot_data_reads_reference_style_mat_files(tmp_path, monkeypatch):
    """A MATLAB file laid out as the reference writes it (variable named by
    the file's path string './shots_time/shot_1.mat', time axis stored as a
    row) loads onto the receivers, the time axis flattened."""
    import scipy.io as sio
    import pysit_amd as pa
    monkeypatch.chdir(tmp_path)
    m, shots = _c1_shots_with_data(1, seed=3)
    os.makedirs("shots_time")
    data = np.arange(50 * 91, dtype=np.float64).reshape(50, 91) * 0.5
    ts = np.arange(50) * 0.002
    sio.savemat("./shots_time/shot_1.mat", mdict={"./shots_time/shot_1.mat": data})
    sio.savemat("./shots_time/ts_1.mat", mdict={"./shots_time/ts_1.mat": ts})
    assert sio.loadmat("./shots_time/ts_1.mat")["./shots_time/ts_1.mat"].shape == (1, 50)
    pa.load_shot_data(shots, "savemat")
    assert np.array_equal(shots[0].receivers.ts, ts)
    assert np.array_equal(shots[0].receivers.data, data)


def test_shot_data_file_errors(tmp_path, monkeypatch):
    """Missing directory, unknown method, h5py absent, a file from another
    method, and a pickle that names anything but numpy arrays: each raises
    instead of loading."""
    import pickle
    import pysit_amd as pa
    monkeypatch.chdir(tmp_path)
    m, shots = _c1_shots_with_data(2)
    with pytest.raises(ImportError, match="no shot directory"):
        pa.load_shot_data(shots, "npy")
    with pytest.raises(TypeError, match="Unknown save_method"):
        pa.save_shot_data(shots, "segy")
    try:
        import h5py  # noqa: F401
    except ImportError:
        with pytest.raises(ImportError, match="h5py"):
            pa.save_shot_data(shots, "h5py")
    pa.save_shot_data(shots, "npy")
    with pytest.raises(IOError, match="different save_method"):
        pa.load_shot_data(shots, "savemat")
    pa.save_shot_data(shots, "pickle")
    ts2 = np.arange(50) * 0.003
    with open("shots_time/ts_1.save", "wb") as fh:
        pickle.dump(ts2, fh, protocol=2)  # the older protocol's array pickle loads too
    pa.load_shot_data(shots, "pickle")
    assert np.array_equal(shots[0].receivers.ts, ts2)
    with open("shots_time/ts_2.save", "wb") as fh:
        pickle.dump(math.sqrt, fh)  # a global other than numpy's array rebuild
    with pytest.raises(IOError, match="different save_method"):
        pa.load_shot_data(shots, "pickle")


def test_batch_limits_follow_the_kernels_addressable_region():
    """A batch's per-node arrays are reached through 32-bit buffer offsets:
    the host's limit is the kernels' kMaxRsrcBytes, and the default batch
    split keeps every batch within it (C4's eight 288^3 shots, C2's 16
    shots: one batch each)."""
    from pysit_amd import device
    src = open(os.path.join(ROOT, "pysit_amd", "csrc", "common.h")).read()
    assert int(re.search(r"kMaxRsrcBytes = (0x[0-9a-fA-F]+)LL", src).group(1), 16) == device.MAX_REGION_BYTES
    lim = device.MAX_REGION_BYTES
    n3 = device.shot_nodes((288, 288, 288), 4)
    assert n3 == 288 ** 3
    k3 = device.shots_per_batch((288, 288, 288), 4)
    assert k3 == 11 and k3 * n3 * 8 <= lim < (k3 + 1) * n3 * 8   # FP64 accumulator
    n2 = device.shot_nodes((1041, 341), 4)
    assert n2 == 1041 * 352                                         # rows at 128 B
    k2 = device.shots_per_batch((1041, 341), 4)
    assert k2 * n2 * 28 <= lim < (k2 + 1) * n2 * 28                 # seven FP32 state buffers
    k2d = device.shots_per_batch((1041, 341), 8)
    assert k2d * n2 * 56 <= lim < (k2d + 1) * n2 * 56
    assert device.shots_per_batch((2000, 2000, 2000), 4) == 1


def test_binomial_schedule_is_optimal():
    """The binomial checkpoint schedule's split (device.binomial_split)
    reverses n steps with s free snapshots in the fewest forward steps: equal
    to a brute-force search over every split, and to the closed form
    t n - C(S + t, S + 1) with S = s + 1 and t the least with C(S + t, S) >= n
    (Griewank and Walther)."""
    from functools import lru_cache
    from math import comb
    from pysit_amd.device import binomial_cost

    @lru_cache(None)
    def brute(n, s):
        if n <= 1:
            return 0
        if s == 0:
            return n * (n - 1) // 2
        return min(m + brute(n - m, s - 1) + brute(m, s) for m in range(1, n))

    for s in range(0, 6):
        for n in range(1, 120):
            assert binomial_cost(n, s) == brute(n, s), (n, s)
            if s and n > 1:
                S, t = s + 1, 0
                while comb(S + t, S) < n:
                    t += 1
                assert binomial_cost(n, s) == t * n - comb(S + t, S + 1)


def test_psi_tile_flag_matches_a_node_by_node_count():
    """3D tile flags: bit2 (no psi) is set exactly for the tiles in whose
    reach (owned nodes +- order/2) no node lies in two or more PML bands,
    counted node by node; bit0 (PML path) whenever a node of the reach lies
    in a PML band or on the grid's edge."""
    from types import SimpleNamespace
    from pysit_amd.device import DeviceBatch, z_tiles_2d
    shape = (70, 41, 150)
    widths = ((5, 6), (4, 0), (7, 3))
    kpml = [(np.full(l, 10.0), np.full(r, 10.0)) for l, r in widths]
    fake = object.__new__(DeviceBatch)
    fake.R, fake.dim, fake.tile_x, fake.tile_y, fake.zalign = 4, 3, 16, 8, 32
    (ntx, nty, ntz), flags = fake._tiles_of(SimpleNamespace(shape=shape, kpml=kpml))
    band = [np.zeros(n, dtype=int) for n in shape]
    edge = [np.zeros(n, dtype=bool) for n in shape]
    for a, (n, (l, r)) in enumerate(zip(shape, widths)):
        band[a][:l] = 1
        if r:
            band[a][n - r:] = 1
        edge[a][[0, n - 1]] = True
    count = band[0][:, None, None] + band[1][None, :, None] + band[2][None, None, :]
    onedge = edge[0][:, None, None] | edge[1][None, :, None] | edge[2][None, None, :]
    zt = z_tiles_2d(shape[2], 32)
    R = 4
    for tx in range(ntx):
        xs = slice(max(0, tx * 16 - R), min(shape[0], (tx + 1) * 16 + R))
        for ty in range(nty):
            ys = slice(max(0, ty * 8 - R), min(shape[1], (ty + 1) * 8 + R))
            for tz, (z0, zhi) in enumerate(zt):
                zs = slice(max(0, z0 - R), min(shape[2], zhi + R))
                f = int(flags[(tx * nty + ty) * ntz + tz])
                c = count[xs, ys, zs]
                assert bool(f & 4) == (c.max() <= 1), (tx, ty, tz)
                assert bool(f & 1) == bool(c.max() >= 1 or onedge[xs, ys, zs].any()), (tx, ty, tz)
    assert (flags & 4).sum() and ((flags & 1) & ~(flags >> 2) & 1).sum()
