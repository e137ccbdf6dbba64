"""Synthetic propagation problems shared by the tests.

Each problem is built twice from the same numbers: once as the oracle's
``Problem`` (flat entry lists) and once as the device ``ShotGeometry``.
The PML profiles follow the reference's closed form (pysit/core/domain.py:
236-254 with mesh.py:749-782): left sigma[j] = A((n-j)/n)^2, right
sigma[j] = A((j+1)/n)^2.
"""
import numpy as np


def pml_pair(n, amp):
    if n == 0:
        return np.zeros(0), np.zeros(0)
    s = amp * np.linspace(0.0, 1.0, n + 1) ** 2
    return s[::-1][:-1].copy(), s[1:].copy()


def mesh_sigma(n_axis, left, right):
    sig = np.zeros(n_axis)
    if left.size:
        sig[:left.size] += left
    if right.size:
        sig[n_axis - right.size:] += right
    return sig


def gaussian_patch(center, shape, width, rng, dim):
    """Entries of a (2w+1)^dim patch around center clipped to the grid."""
    rngs = [np.arange(max(0, c - width), min(n, c + width + 1)) for c, n in zip(center, shape)]
    grid = np.meshgrid(*rngs, indexing="ij")
    pos = np.stack([g.reshape(-1) for g in grid], axis=1)
    d2 = sum((pos[:, i] - center[i]) ** 2 for i in range(dim))
    w = np.exp(-0.5 * d2 / max(1.0, width / 2) ** 2) * (1.0 + 0.1 * rng.random(pos.shape[0]))
    return pos, w


class Synthetic:
    """One shot on a padded grid with PML of widths `pml` per axis side."""

    def __init__(self, shape, pml, deltas, nt, order, seed=0, amp_pml=60.0, n_src=1, n_rec=7,
                 window=None, model=None, dt=None, src_center=None, rec_z=None):
        rng = np.random.default_rng(seed)
        self.shape = tuple(shape)
        dim = len(shape)
        self.dim = dim
        self.deltas = tuple(deltas)
        self.order = order
        self.nt = nt
        self.pml = pml  # per axis (nl, nr)
        self.kpml = [pml_pair(nl, amp_pml) if nl else (np.zeros(0), np.zeros(0)) for nl, nr in pml]
        # allow different right width
        self.kpml = []
        for (nl, nr) in pml:
            l = pml_pair(nl, amp_pml)[0] if nl else np.zeros(0)
            r = pml_pair(nr, amp_pml * 1.3)[1] if nr else np.zeros(0)
            self.kpml.append((l, r))
        self.msig = [mesh_sigma(n, l, r) for n, (l, r) in zip(shape, self.kpml)]
        # model: unpadded interior sizes, edge padded by clamping
        inner = tuple(n - nl - nr for n, (nl, nr) in zip(shape, pml))
        if model is None:
            base = 1.0 + 0.3 * rng.random(inner)
            model = base
        self.model = np.asarray(model, dtype=np.float64)
        if window is None:
            window = [(nl, 0, n_in - 1) for (nl, nr), n_in in zip(pml, inner)]
        self.window = window
        self.Cpad = self._pad_model()
        vmax = self.Cpad.max()
        self.dt = dt if dt is not None else (1.0 / 6.0) * min(deltas) / vmax
        # sources
        pos_all, chan_all, w_all = [], [], []
        for c in range(n_src):
            center = [int(rng.integers(pml[i][0] + 2, shape[i] - pml[i][1] - 2)) for i in range(dim)]
            if src_center is not None:
                center = list(src_center)
            pos, w = gaussian_patch(center, shape, 2, rng, dim)
            pos_all.append(pos)
            chan_all.append(np.full(pos.shape[0], c))
            w_all.append(w)
        # the reference's point weights integrate to one: ~1/prod(deltas)
        self.src = (np.concatenate(pos_all), np.concatenate(chan_all),
                    np.concatenate(w_all) / float(np.prod(deltas)))
        t = np.arange(nt + 1) * self.dt
        f0 = 1.0 / (12 * self.dt)
        amp = np.zeros((nt + 1, n_src))
        for c in range(n_src):
            tau = t - (1.5 + 0.3 * c) / f0
            x = (np.pi * f0 * tau) ** 2
            amp[:, c] = (1 - 2 * x) * np.exp(-x)
        self.amp = amp
        # receivers: a line along x (2D) or a patch (3D) near the top
        rec_pos, rec_chan, rec_w = [], [], []
        zr = pml[-1][0] + 2 if rec_z is None else int(rec_z)
        for r in range(n_rec):
            if dim == 2:
                xr = int(np.linspace(pml[0][0] + 1, shape[0] - pml[0][1] - 2, n_rec)[r])
                center = [xr, zr]
            else:
                xr = int(np.linspace(pml[0][0] + 1, shape[0] - pml[0][1] - 2, n_rec)[r])
                yr = int(rng.integers(pml[1][0] + 1, shape[1] - pml[1][1] - 1))
                center = [xr, yr, zr]
            pos, w = gaussian_patch(center, shape, 1, rng, dim)
            rec_pos.append(pos)
            rec_chan.append(np.full(pos.shape[0], r))
            rec_w.append(w)
        rp = np.concatenate(rec_pos)
        self.rec_pos = rp
        self.rec_chan = np.concatenate(rec_chan)
        self.rec_w = np.concatenate(rec_w)
        self.n_rec = n_rec
        self.rec_idx = np.ravel_multi_index(tuple(rp.T), shape)

    def _pad_model(self):
        idx = []
        for i, (off, lo, hi) in enumerate(self.window):
            idx.append(np.clip(np.arange(self.shape[i]) - off, lo, hi))
        return self.model[np.ix_(*idx)]

    def flat(self, pos):
        return np.ravel_multi_index(tuple(np.asarray(pos).T), self.shape)

    def oracle_problem(self):
        from oracle.oracle import Problem
        src = (self.flat(self.src[0]), self.src[1], self.src[2])
        rec = (self.rec_idx, self.rec_chan, self.rec_w * 0.5)
        adj = (self.rec_idx, self.rec_chan, self.rec_w)
        return Problem(self.shape, self.deltas, self.dt, self.nt, self.Cpad, self.kpml, self.msig,
                       src, self.amp, rec, self.n_rec, adj=adj, order=self.order, pad=self.pml)

    def geometry(self):
        from pysit_amd.device import ShotGeometry
        return ShotGeometry(self.shape, self.kpml, self.window, self.nt, self.src, self.amp,
                            (self.rec_idx, self.rec_chan, self.rec_w * 0.5), self.n_rec,
                            adj=(self.rec_pos, self.rec_chan, self.rec_w))


def oracle_from_solver(solver, shot, order=None):
    """The oracle Problem of one shot, built from the solver's host setup:
    edge-padded C, the kernel PML arrays, the mesh sigma per axis and the
    point-operator entries (flat padded indices)."""
    from oracle.oracle import Problem
    from pysit_amd.core.receivers import adjoint_entries, sampling_entries
    mesh = solver.mesh
    shape = tuple(mesh.shape(include_bc=True, as_grid=True))
    oc = solver.operator_components
    if mesh.dim == 2:
        msig = [oc.sx.reshape(shape)[:, 0], oc.sz.reshape(shape)[0, :]]
    else:
        msig = [oc.sx.reshape(shape)[:, 0, 0], oc.sy.reshape(shape)[0, :, 0], oc.sz.reshape(shape)[0, 0, :]]
    (spos, schan, sw), amp = shot.sources.injection(solver.dt, solver.nsteps)
    src = (np.ravel_multi_index(tuple(spos.T), shape), schan, sw)
    rec = sampling_entries(shot.receivers, shape)
    apos, achan, aw = adjoint_entries(shot.receivers)
    adj = (np.ravel_multi_index(tuple(apos.T), shape), achan, aw)
    return Problem(shape, mesh.deltas, solver.dt, solver.nsteps, np.asarray(solver.model_parameters.C, dtype=np.float64),
                   solver.kernel_pml(), msig, src, amp, rec, shot.receivers.receiver_count, adj=adj,
                   order=order or solver.spatial_accuracy_order, pad=mesh.pad_widths())


def c1_setup(order=4, precision="double", tf=1.5, nx=91, nz=71, pml=0.1):
    """The horizontal-reflector case: x in [0, 0.9], z in [0, 0.7], dx = 0.01,
    v = 1 + 0.5 exp(-((z - 0.45)/0.02)^2), PML(0.1, 100) on all sides, one
    Ricker (10 Hz) source at depth 0.05 mid-x, receivers at every x node at
    depth 0.05."""
    import pysit_amd as pa
    px = pa.PML(pml, 100) if pml else pa.Dirichlet()
    d = pa.RectangularDomain((0.0, 0.9, px, px), (0.0, 0.7, px, px))
    m = pa.CartesianMesh(d, nx, nz)
    XX, ZZ = m.mesh_coords()
    C = (1.0 + 0.5 * np.exp(-((ZZ - 0.45) / 0.02) ** 2)).reshape(-1, 1)
    shots = pa.equispaced_acquisition(m, pa.RickerWavelet(10.0), sources=1, source_depth=0.05,
                                      receivers="max", receiver_depth=0.05)
    solver = pa.ConstantDensityAcousticWave(m, spatial_accuracy_order=order, trange=(0.0, tf),
                                            kernel_implementation="hip", precision=precision)
    model = solver.ModelParameters(m, {"C": C})
    return m, shots, solver, model


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def oracle_from_geometry(g, model, deltas, dt, order):
    """The oracle Problem of one device ShotGeometry over ``model`` (the
    unpadded wave speed the geometry's window reads): the edge-padded C by
    the window's clamping, the mesh sigma laid out from the kernel arrays
    (left by node, right on the last nodes), flat entry lists, the adjoint
    entries in padded coordinates."""
    from oracle.oracle import Problem
    model = np.asarray(model, dtype=np.float64)
    idx = [np.clip(np.arange(n) - off, lo, hi) for n, (off, lo, hi) in zip(g.shape, g.window)]
    Cpad = model[np.ix_(*idx)]
    msig = [mesh_sigma(n, l, r) for n, (l, r) in zip(g.shape, g.kpml)]
    pos, chan, w = g.src
    src = (np.ravel_multi_index(tuple(pos.T), g.shape), chan, w)
    apos, achan, aw = g.adj
    adj = (np.ravel_multi_index(tuple(apos.T), g.shape), achan, aw)
    pads = []
    for n, (off, lo, hi) in zip(g.shape, g.window):
        left = off + lo
        pads.append((left, n - left - (hi - lo + 1)))
    return Problem(g.shape, deltas, dt, g.nt, Cpad, g.kpml, msig, src, g.amp[:g.nt], g.rec, g.nrec, adj=adj,
                   order=order, pad=pads)


def truncated_geometry(g, nt):
    """The same shot with only its first ``nt`` steps."""
    from pysit_amd.device import ShotGeometry
    return ShotGeometry(g.shape, g.kpml, g.window, nt, g.src, g.amp[:nt], g.rec, g.nrec, adj=g.adj)


def pml_golden_case(tag, precision="double"):
    """The cases of tests/golden/pml_golden.npz built with this package's
    API (tests/golden/make_pml_golden.py builds them with PySIT's): "c1" and
    "c3d" at spatial order 4, "c1o8" and "c3do8" at order 8; returns (mesh,
    shots, solver, true C, start C), C as (dof, 1) arrays."""
    import pysit_amd as pa
    order = 8 if tag.endswith("o8") else 4
    tag = tag[:-2] if tag.endswith("o8") else tag
    if tag == "c1":
        p = pa.PML(0.1, 100)
        d = pa.RectangularDomain((0.0, 0.9, p, p), (0.0, 0.7, p, p))
        m = pa.CartesianMesh(d, 91, 71)
        X, Z = m.mesh_coords()
        C = (1.0 + 0.5 * np.exp(-((Z - 0.45) / 0.02) ** 2)).reshape(-1, 1)
        C0 = (1.0 + 0.3 * np.exp(-((Z - 0.45) / 0.02) ** 2)).reshape(-1, 1)
        shots = pa.equispaced_acquisition(m, pa.RickerWavelet(10.0), sources=1, source_depth=0.05,
                                          receivers="max", receiver_depth=0.05)
        tr = (0.0, 1.5)
    else:
        p = pa.PML(0.08, 60)
        d = pa.RectangularDomain((0.0, 0.3, p, p), (0.0, 0.3, p, p), (0.0, 0.3, p, p))
        m = pa.CartesianMesh(d, 16, 16, 16)
        X, Y, Z = m.mesh_coords()
        C = (1.0 + 0.4 * (Z > 0.17) + 0.15 * np.exp(-((X - 0.15) ** 2 + (Y - 0.15) ** 2 + (Z - 0.1) ** 2) / 0.004))
        C0 = 1.0 + 0.4 * (Z > 0.17)
        C, C0 = C.reshape(-1, 1), C0.reshape(-1, 1)
        rec = np.linspace(0.05, 0.25, 4)
        src = pa.PointSource(m, (0.15, 0.15, 0.04), pa.RickerWavelet(12.0))
        recs = pa.ReceiverSet(m, [pa.PointReceiver(m, (a, b, 0.04)) for a in rec for b in rec])
        shots = [pa.Shot(src, recs)]
        tr = (0.0, 0.25)
    solver = pa.ConstantDensityAcousticWave(m, spatial_accuracy_order=order, trange=tr, kernel_implementation="hip",
                                            precision=precision)
    return m, shots, solver, C, C0


def golden_operand(nt, nrec):
    """The closed-form migrate operand of tests/golden/make_pml_golden.py."""
    k = np.arange(nt)[:, None]
    j = np.arange(nrec)[None, :]
    return np.sin(0.37 * k + 1.3 * j) * np.exp(-k / 800.0)


def golden_direction(m):
    """The closed-form model perturbation of tests/golden/make_pml_golden.py
    ((dof, 1), physical grid)."""
    if m.dim == 3:
        X, Y, Z = m.mesh_coords()
        return (0.02 * np.sin(7.0 * X + 1.0) * np.cos(5.0 * Z) * np.cos(3.0 * Y)).reshape(-1, 1)
    X, Z = m.mesh_coords()
    return (0.02 * np.sin(7.0 * X + 1.0) * np.cos(5.0 * Z)).reshape(-1, 1)
