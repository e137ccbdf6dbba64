"""Synthetic benchmark and test configurations.

Each builder returns ``(mesh, shots, solver, model)`` (plus extras) in the
package's own API, seeded so that every run sees the same numbers:

* ``horizontal_reflector_2d`` -- the reference's CPU case: 91 x 71 nodes on
  [0, 0.9] x [0, 0.7], v = 1 + 0.5 exp(-((z-0.45)/0.02)^2), PML(0.1, 100),
  one 10 Hz Ricker at depth 0.05, 91 receivers, T = 1.5.
* ``layered_2d`` -- nz x nx with spacing dx, velocities 1500..vmax in
  ``n_layers`` dipping layers (seeded dips), times (1 + 0.05 g) with g a
  Gaussian random field smoothed with sigma = 150 m and scaled to unit RMS,
  clipped to [vmin, vmax]; PML of ``pml_cells`` nodes.
* ``uneven_2d`` -- shots of seeded random aperture (a window of the layered
  model's columns) and step count, as device geometries for one batch.
* ``layered_3d`` -- the 3D analogue with a Gaussian low-velocity anomaly.
* ``fwi_2d`` / ``multiband_fwi`` -- the 2D FWI case (layered true model,
  smoothed start model) and its frequency-continuation LBFGS loop.
"""
import numpy as np
from scipy.ndimage import gaussian_filter

from .core import (RectangularDomain, PML, Dirichlet, CartesianMesh, RickerWavelet, PointSource, PointReceiver,
                   ReceiverSet, Shot, equispaced_acquisition)
from .solvers import ConstantDensityAcousticWave

__all__ = ["horizontal_reflector_2d", "layered_velocity_2d", "layered_2d", "uneven_2d", "layered_velocity_3d",
           "layered_3d", "fwi_2d", "multiband_fwi", "model_error", "pml_amplitude", "trange_for_steps"]


def pml_amplitude(vmax, width, reflection=1e-3):
    """Quadratic-profile damping amplitude for a target normal-incidence
    reflection: 3 vmax ln(1/R) / (2 width)."""
    return 3.0 * vmax * np.log(1.0 / reflection) / (2.0 * width)


def trange_for_steps(dt, nt):
    """A time range whose ceil((tf - t0)/dt) is exactly nt."""
    return (0.0, (nt - 0.5) * dt)


def horizontal_reflector_2d(order=4, precision="double", tf=1.5):
    p = PML(0.1, 100)
    d = RectangularDomain((0.0, 0.9, p, p), (0.0, 0.7, p, p))
    m = CartesianMesh(d, 91, 71)
    X, Z = m.mesh_coords()
    C = (1.0 + 0.5 * np.exp(-((Z - 0.45) / 0.02) ** 2)).reshape(-1, 1)
    shots = equispaced_acquisition(m, RickerWavelet(10.0), sources=1, source_depth=0.05, receivers="max",
                                   receiver_depth=0.05)
    solver = ConstantDensityAcousticWave(m, spatial_accuracy_order=order, trange=(0.0, tf),
                                         kernel_implementation="hip", precision=precision)
    return m, shots, solver, solver.ModelParameters(m, {"C": C})


def layered_velocity_2d(nx, nz, dx, n_layers=12, vmin=1500.0, vmax=4500.0, seed=0, field_sigma=150.0,
                        field_amp=0.05):
    """(nx, nz) velocity: dipping layers plus a smoothed random field."""
    rng = np.random.default_rng(seed)
    x = np.arange(nx) * dx
    z = np.arange(nz) * dx
    depth = z[-1]
    tops = np.sort(rng.uniform(0.05 * depth, 0.95 * depth, n_layers - 1))
    dips = rng.uniform(-0.12, 0.12, n_layers - 1)  # slopes (about +-7 degrees)
    xc = x.mean()
    vel_layers = np.linspace(vmin, vmax, n_layers)
    v = np.full((nx, nz), vel_layers[0])
    for i in range(n_layers - 1):
        iface = tops[i] + dips[i] * (x - xc)
        below = z[None, :] >= iface[:, None]
        v[below] = vel_layers[i + 1]
    g = gaussian_filter(rng.standard_normal((nx, nz)), sigma=field_sigma / dx, mode="reflect")
    g /= np.sqrt(np.mean(g ** 2))
    v = v * (1.0 + field_amp * g)
    return np.clip(v, vmin, vmax)


def layered_2d(nx=1001, nz=301, dx=10.0, n_shots=16, nt=3000, order=8, precision="single", pml_cells=20,
               f0=8.0, depth=20.0, seed=0, n_receivers=None, model=None):
    """The 2D layered case: ``n_shots`` Ricker sources evenly spaced at
    ``depth``, one receiver per surface node at ``depth``, ``nt`` steps at the
    CFL dt of the model."""
    v = layered_velocity_2d(nx, nz, dx, seed=seed) if model is None else model
    width = pml_cells * dx
    A = pml_amplitude(v.max(), width)
    p = PML(width, A)
    d = RectangularDomain((0.0, (nx - 1) * dx, p, p), (0.0, (nz - 1) * dx, p, p))
    m = CartesianMesh(d, nx, nz)
    shots = equispaced_acquisition(m, RickerWavelet(f0), sources=n_shots, source_depth=depth,
                                   receivers=n_receivers or nx, receiver_depth=depth)
    dt = (1.0 / 6.0) * dx / v.max()
    solver = ConstantDensityAcousticWave(m, spatial_accuracy_order=order, trange=trange_for_steps(dt, nt),
                                         kernel_implementation="hip", precision=precision)
    model = solver.ModelParameters(m, {"C": v.reshape(-1, 1)})
    solver.model_parameters = model
    assert solver.nsteps == nt, (solver.nsteps, nt)
    return m, shots, solver, model


def uneven_2d(n_shots=64, nx=1001, nz=301, dx=10.0, order=8, pml_cells=20, f0=8.0, depth=20.0,
              aperture=(300, 1001), nt_range=(1500, 4000), seed=0, model=None):
    """Shots of uneven aperture and recording length over the layered model.

    Shot s sees columns [a, a + w) of the nx x nz model, w and the step count
    drawn uniformly from ``aperture`` and ``nt_range`` (seeded); its padded
    grid is (w + 2 pml_cells) x (nz + 2 pml_cells), edge-padded from that
    window.  One Ricker source at the aperture centre, one receiver per
    aperture column, both at ``depth``.  All shots share the CFL dt of the
    whole model so they run in one batch.

    Returns (geometries, model (nx, nz), deltas, dt, per-shot metadata)."""
    from .core.receivers import sampling_entries, adjoint_entries
    from .device import ShotGeometry
    v = layered_velocity_2d(nx, nz, dx, seed=seed) if model is None else np.asarray(model, dtype=np.float64)
    rng = np.random.default_rng(seed + 1)
    width = pml_cells * dx
    A = pml_amplitude(v.max(), width)
    dt = (1.0 / 6.0) * dx / v.max()
    wav = RickerWavelet(f0)
    geos, meta = [], []
    for _ in range(n_shots):
        w = int(rng.integers(aperture[0], aperture[1] + 1))
        a = int(rng.integers(0, nx - w + 1))
        nt = int(rng.integers(nt_range[0], nt_range[1] + 1))
        p = PML(width, A)
        d = RectangularDomain((a * dx, (a + w - 1) * dx, p, p), (0.0, (nz - 1) * dx, p, p))
        m = CartesianMesh(d, w, nz)
        src = PointSource(m, ((a + 0.5 * (w - 1)) * dx, depth), wav)
        recs = ReceiverSet(m, [PointReceiver(m, ((a + i) * dx, depth)) for i in range(w)])
        shape = tuple(m.shape(include_bc=True, as_grid=True))
        inj, amp = src.injection(dt, nt)
        kpml = []
        for prm in (m.parameters[0], m.parameters[1]):
            kpml.append((prm.lbc.sigma if prm.lbc.type == "pml" else np.zeros(0),
                         prm.rbc.sigma if prm.rbc.type == "pml" else np.zeros(0)))
        window = [(pml_cells - a, a, a + w - 1), (pml_cells, 0, nz - 1)]
        geos.append(ShotGeometry(shape, kpml, window, nt, inj, amp, sampling_entries(recs, shape), w,
                                 adj=adjoint_entries(recs)))
        meta.append({"a": a, "w": w, "nt": nt, "shape": shape})
    return geos, v, (dx, dx), dt, meta


def layered_velocity_3d(n, dx, vmin=1500.0, vmax=4000.0, n_layers=8, seed=0, anomaly_radius=400.0,
                        anomaly=-0.10):
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    z = np.arange(nz) * dx
    tops = np.sort(rng.uniform(0.05 * z[-1], 0.95 * z[-1], n_layers - 1))
    vel = np.linspace(vmin, vmax, n_layers)
    v1 = np.full(nz, vel[0])
    for i in range(n_layers - 1):
        v1[z >= tops[i]] = vel[i + 1]
    v = np.broadcast_to(v1[None, None, :], (nx, ny, nz)).copy()
    X, Y, Z = np.meshgrid(np.arange(nx) * dx, np.arange(ny) * dx, z, indexing="ij")
    c = (X.mean(), Y.mean(), Z.mean())
    r2 = (X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2
    v *= 1.0 + anomaly * np.exp(-r2 / anomaly_radius ** 2)
    return np.clip(v, vmin, vmax)


def layered_3d(n=256, dx=12.5, n_shots=8, nt=2000, order=8, precision="single", pml_cells=16, f0=6.0,
               rec_patch=64, depth=None, seed=0, pml_reflection=1e-3, unit=1e-3):
    """The 3D case: n^3 nodes, ``n_shots`` sources in a line at the surface,
    a rec_patch x rec_patch receiver grid at the surface.  Lengths and
    speeds are given in metres and m/s and scaled by ``unit``; times stay in
    seconds.  The default, km and km/s, matters: the 3D step multiplies the
    sigma_x sigma_y sigma_z psi term by C^2
    (constant_density_acoustic_time_scalar_3D.h:218), so its PML depends on
    the unit of speed.  With C near 4 km/s it absorbs; with C near 4000 m/s
    the PML corners grow without bound within a few hundred steps (the
    reference's own arithmetic, reproduced by the oracle and the kernels)."""
    v = layered_velocity_3d((n, n, n), dx, seed=seed) * unit
    dx = dx * unit
    width = pml_cells * dx
    A = pml_amplitude(v.max(), width, pml_reflection)
    p = PML(width, A)
    L = (n - 1) * dx
    d = RectangularDomain((0.0, L, p, p), (0.0, L, p, p), (0.0, L, p, p))
    m = CartesianMesh(d, n, n, n)
    depth = 2 * dx if depth is None else depth
    wav = RickerWavelet(f0)
    xs = np.linspace(0.2 * L, 0.8 * L, n_shots)
    rx = np.linspace(0.0, L, rec_patch)
    shots = []
    for x in xs:
        recs = ReceiverSet(m, [PointReceiver(m, (a, b, depth)) for a in rx for b in rx])
        shots.append(Shot(PointSource(m, (x, L / 2, depth), wav), recs))
    dt = (1.0 / 6.0) * dx / v.max()
    solver = ConstantDensityAcousticWave(m, spatial_accuracy_order=order, trange=trange_for_steps(dt, nt),
                                         kernel_implementation="hip", precision=precision)
    model = solver.ModelParameters(m, {"C": v.reshape(-1, 1)})
    solver.model_parameters = model
    return m, shots, solver, model


def fwi_2d(nx=601, nz=201, dx=15.0, nt=2500, order=8, precision="single", pml_cells=20, seed=0,
           start_sigma=300.0):
    """The 2D FWI case: the layered model family on nx x nz nodes as the true
    model, and the true model Gaussian-smoothed with ``start_sigma`` metres
    (edges extended) as the start model.  The time range gives ``nt`` steps
    at the CFL dt of the true model; every other model the solver takes
    re-derives dt (and the step count) from its own maximum speed.

    Returns (mesh, solver, true ModelParameters, start ModelParameters)."""
    v = layered_velocity_2d(nx, nz, dx, seed=seed)
    v0 = gaussian_filter(v, sigma=start_sigma / dx, mode="nearest")
    width = pml_cells * dx
    p = PML(width, pml_amplitude(v.max(), width))
    d = RectangularDomain((0.0, (nx - 1) * dx, p, p), (0.0, (nz - 1) * dx, p, p))
    m = CartesianMesh(d, nx, nz)
    dt = (1.0 / 6.0) * dx / v.max()
    solver = ConstantDensityAcousticWave(m, spatial_accuracy_order=order, trange=trange_for_steps(dt, nt),
                                         kernel_implementation="hip", precision=precision)
    true_model = solver.ModelParameters(m, {"C": v.reshape(-1, 1)})
    start_model = solver.ModelParameters(m, {"C": v0.reshape(-1, 1)})
    solver.model_parameters = true_model
    assert solver.nsteps == nt, (solver.nsteps, nt)
    return m, solver, true_model, start_model


class _CountedObjective(object):
    """TemporalLeastSquares that counts its evaluations, gradients and
    algorithmic grid-point updates (forward: shots x padded nodes x steps at
    the model's dt; gradient: twice that)."""

    def __init__(self, solver, n_nodes, inner=None, **kwargs):
        from .objective_functions import TemporalLeastSquares
        self.inner = inner if inner is not None else TemporalLeastSquares(solver, **kwargs)
        self.solver = solver
        self.n_nodes = n_nodes
        self.evaluations = self.gradients = 0
        self.updates = 0.0

    def use_parallel(self):
        return self.inner.use_parallel()

    @property
    def parallel_wrap_shot(self):
        return self.inner.parallel_wrap_shot

    def evaluate(self, shots, m0, **kwargs):
        v = self.inner.evaluate(shots, m0, **kwargs)
        self.evaluations += 1
        self.updates += len(shots) * self.n_nodes * self.solver.nsteps
        return v

    def compute_gradient(self, shots, m0, aux_info={}, **kwargs):
        g = self.inner.compute_gradient(shots, m0, aux_info=aux_info, **kwargs)
        self.gradients += 1
        self.updates += 2.0 * len(shots) * self.n_nodes * self.solver.nsteps
        return g


def model_error(model, true_model, depth_fraction=None):
    """Relative L2 error of the wave speed over the unpadded domain, or over
    its upper ``depth_fraction`` of depth (the last axis)."""
    c, ct = np.asarray(model.C), np.asarray(true_model.C)
    if depth_fraction is not None:
        shape = model.mesh.shape(as_grid=True)
        nz = shape[-1]
        keep = max(1, int(round(depth_fraction * nz)))
        c = c.reshape(shape)[..., :keep]
        ct = ct.reshape(shape)[..., :keep]
    c, ct = c.ravel(), ct.ravel()
    return float(np.linalg.norm(c - ct) / np.linalg.norm(ct))


def multiband_fwi(mesh, solver, true_model, start_model, bands=(3.0, 5.0, 8.0), iterations=20,
                  memory_length=10, n_shots=32, depth=None, verbose=False, sync=None, log=None,
                  objective=None, generate=None, shallow=1.0 / 3.0):
    """Frequency-continuation FWI with LBFGS over TemporalLeastSquares.

    For each Ricker peak frequency in ``bands`` (low to high): new shots with
    that wavelet (``n_shots`` sources and one receiver per surface node at
    ``depth``, equispaced acquisition), observed data generated in the true
    model (the time path of generate_seismic_data, on the true model's time
    axis), then ``iterations`` LBFGS iterations with ``memory_length`` pairs
    and the backtracking line search, starting from the previous band's
    result.  Per band this is the inversion pattern of
    examples/Marmousi2D.py:18-69 (acquisition, generate_seismic_data,
    TemporalLeastSquares, LBFGS, invalg(shots, initial_value, nsteps,
    line_search='backtrack')); the trial models' dt changes with their maximum
    speed, and the objective interpolates the observed data onto it
    (objective_functions/temporal_least_squares.py:56).

    ``sync`` (optional) is called before each band's timer is read, e.g. a
    device synchronisation; ``log`` (a file) takes the LBFGS progress lines
    when ``verbose``.  ``objective(solver)`` and ``generate(shots, solver,
    model)`` replace TemporalLeastSquares and generate_seismic_data (another
    implementation of the same objective, run through the same loop).
    Returns (final model, one report dict per band: frequency, objective
    history, misfit reduction, model error before and after over the whole
    model and over its upper ``shallow`` fraction of depth, evaluations,
    gradients, grid-point updates, seconds)."""
    import time
    from .modeling import generate_seismic_data
    from .optimization import LBFGS
    depth = 2 * mesh.x.delta if depth is None else depth
    n_nodes = int(np.prod(mesh.shape(include_bc=True, as_grid=True)))
    model = start_model
    reports = []
    for f0 in bands:
        t0 = time.perf_counter()
        shots = equispaced_acquisition(mesh, RickerWavelet(f0), sources=n_shots, source_depth=depth,
                                       receivers="max", receiver_depth=depth)
        (generate or generate_seismic_data)(shots, solver, true_model)
        obj = _CountedObjective(solver, n_nodes, inner=objective(solver) if objective is not None else None)
        obj.updates += len(shots) * n_nodes * solver.nsteps
        err0 = model_error(model, true_model)
        err0_shallow = model_error(model, true_model, shallow)
        invalg = LBFGS(obj, memory_length=memory_length)
        if log is not None:
            invalg.logfile = log
        status = {"objective_frequency": 1, "gradient_length_frequency": 1, "alpha_frequency": 1}
        model = invalg(shots, model, iterations, line_search="backtrack", status_configuration=status,
                       verbose=verbose)
        # objective of the final model (the history holds the iterates before each step)
        f_final = obj.evaluate(shots, model)
        if sync is not None:
            sync()
        hist = [float(v) for _, v in sorted(invalg.objective_history.items())]
        reports.append({
            "f0_hz": float(f0), "iterations": int(iterations), "objective": hist + [float(f_final)],
            "misfit_reduction": float(f_final / hist[0]) if hist and hist[0] > 0 else None,
            "model_error_start": err0, "model_error_final": model_error(model, true_model),
            "shallow_fraction": float(shallow), "model_error_shallow_start": err0_shallow,
            "model_error_shallow_final": model_error(model, true_model, shallow),
            "evaluations": obj.evaluations, "gradients": obj.gradients, "updates": obj.updates,
            "seconds": time.perf_counter() - t0})
    return model, reports
