"""Python side of the CPU oracle (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the CPU result the HIP library is
compared against.  The product package ``pysit_amd`` never imports it.

It loads ``oracle/liborc.so`` (the C restatement in ``cda_oracle.c``) and, when
it was built here from the PySIT sources, ``oracle/_ref/libcda_ref.so`` (the
reference's own step templates).  Besides the marshalling it restates, in
numpy, the small host-side pieces of the reference path that the checks need:

* ``interp_zero_fill`` -- ``scipy.interpolate.interp1d(kind='linear',
  bounds_error=False, fill_value=0.0)`` as used by
  ``pysit/core/receivers.py:75-105`` (pinned in the tests against scipy's
  own interp1d and by closed forms: exact on piecewise-linear data, zero
  outside the range).
* ``misfit`` -- ``0.5 * dt * sum ||r||^2``
  (``pysit/objective_functions/temporal_least_squares.py:67-82``).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_lp = ctypes.POINTER(ctypes.c_long)
_ip = ctypes.POINTER(ctypes.c_int)
_i = ctypes.c_int
_d = ctypes.c_double


def build():
    """Compile the oracle (and the reference shim when the sources exist)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if os.path.isdir("/root/reference/pysit"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _load(path):
    if not os.path.exists(path):
        raise FileNotFoundError(f"oracle library {path} is not built (make -C oracle)")
    return ctypes.CDLL(path)


_LIBS = {}


def lib(fast=False):
    key = "fast" if fast else "exact"
    if key not in _LIBS:
        L = _load(os.path.join(HERE, "liborc_fast.so" if fast else "liborc.so"))
        _declare(L, "orc")
        common = ([_i] * 5 + [_d] * 4 + [_i, _dp] + [_dp, _i] * 6 + [_dp] * 3 +
                  [_i, _lp, _ip, _dp, _i, _dp])
        L.orc_forward.restype = _i
        L.orc_forward.argtypes = common + [_i, _lp, _ip, _dp, _i, _dp, _i, _dp, _dp]
        L.orc_migrate.restype = _i
        L.orc_migrate.argtypes = common + [_i, _lp, _ip, _dp, _i, _dp, _i, _i, _dp]
        L.orc_fold.restype = None
        L.orc_fold.argtypes = [_i] * 10 + [_dp, _dp]
        _LIBS[key] = L
    return _LIBS[key]


def ref_available(fast=False):
    return os.path.exists(os.path.join(HERE, "_ref", "libcda_ref_fast.so" if fast else "libcda_ref.so"))


def ref_lib(fast=False):
    key = "ref_fast" if fast else "ref"
    if key not in _LIBS:
        L = _load(os.path.join(HERE, "_ref", "libcda_ref_fast.so" if fast else "libcda_ref.so"))
        _declare(L, "ref")
        for suf, P in (("d", _dp), ("f", _fp)):
            f2 = getattr(L, f"ref_step2d_omp_{suf}", None)
            if f2 is not None:
                f2.restype = _i
                f2.argtypes = ([_i] + [P] * 6 + [P, _i] * 4 + [_d] * 3 + [_i] * 2 + [P] * 3)
        _LIBS[key] = L
    return _LIBS[key]


def _declare(L, prefix):
    for suf, P in (("d", _dp), ("f", _fp)):
        f2 = getattr(L, f"{prefix}_step2d_{suf}")
        f2.restype = _i
        f2.argtypes = ([_i] + [P] * 6 + [P, _i] * 4 + [_d] * 3 + [_i] * 2 + [P] * 3)
        f3 = getattr(L, f"{prefix}_step3d_{suf}")
        f3.restype = _i
        f3.argtypes = ([_i] + [P] * 8 + [P, _i] * 6 + [_d] * 4 + [_i] * 3 + [P] * 5)


def _ptr(a, P):
    return a.ctypes.data_as(P)


def _arr(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def step2d(order, km1_u, k_phix, k_phiz, k_u, C, rhs, xl, xr, zl, zr, dt, dx, dz, nx, nz,
           dtype=np.float64, use_ref=False, fast=False, omp=False):
    """One 2D step; returns (kp1_phix, kp1_phiz, kp1_u) as flat arrays.
    ``omp`` (with ``use_ref``, orders 4 and 6): the reference's OpenMP
    variant (constant_density_acoustic_time_scalar_2D_{4,6}.h), threaded over
    OMP_NUM_THREADS (set here to the affinity count when unset)."""
    L = ref_lib(fast) if use_ref else lib(fast)
    if omp:
        if not use_ref or order not in (4, 6):
            raise ValueError("the OpenMP variant is the reference's, orders 4 and 6")
        os.environ.setdefault("OMP_NUM_THREADS", str(len(os.sched_getaffinity(0))))
    P = _dp if dtype == np.float64 else _fp
    suf = "d" if dtype == np.float64 else "f"
    ins = [_arr(a, dtype).reshape(-1) for a in (km1_u, k_phix, k_phiz, k_u, C, rhs)]
    pml = [_arr(a, dtype).reshape(-1) for a in (xl, xr, zl, zr)]
    n = nx * nz
    outs = [np.zeros(n, dtype) for _ in range(3)]
    args = [order] + [_ptr(a, P) for a in ins]
    for a in pml:
        args += [_ptr(a, P), a.size]
    args += [dt, dx, dz, nx, nz] + [_ptr(a, P) for a in outs]
    name = ("ref" if use_ref else "orc") + ("_step2d_omp_" if omp else "_step2d_") + suf
    rc = getattr(L, name)(*args)
    if rc != 0:
        raise ValueError(f"step2d failed ({rc})")
    return tuple(outs)


def step3d(order, km1_u, k_phix, k_phiy, k_phiz, k_psi, k_u, C, rhs, xl, xr, yl, yr, zl, zr,
           dt, dx, dy, dz, nx, ny, nz, dtype=np.float64, use_ref=False, fast=False):
    """One 3D step; returns (kp1_phix, kp1_phiy, kp1_phiz, kp1_psi, kp1_u)."""
    L = ref_lib(fast) if use_ref else lib(fast)
    P = _dp if dtype == np.float64 else _fp
    suf = "d" if dtype == np.float64 else "f"
    ins = [_arr(a, dtype).reshape(-1) for a in (km1_u, k_phix, k_phiy, k_phiz, k_psi, k_u, C, rhs)]
    pml = [_arr(a, dtype).reshape(-1) for a in (xl, xr, yl, yr, zl, zr)]
    n = nx * ny * nz
    outs = [np.zeros(n, dtype) for _ in range(5)]
    args = [order] + [_ptr(a, P) for a in ins]
    for a in pml:
        args += [_ptr(a, P), a.size]
    args += [dt, dx, dy, dz, nx, ny, nz] + [_ptr(a, P) for a in outs]
    rc = getattr(L, ("ref" if use_ref else "orc") + f"_step3d_{suf}")(*args)
    if rc != 0:
        raise ValueError(f"step3d failed ({rc})")
    return tuple(outs)


class Problem:
    """Everything one shot propagation needs, as flat FP64 arrays.

    ``shape`` is the padded grid (nx, nz) or (nx, ny, nz); ``deltas`` matches.
    ``kpml`` holds the kernel's PML arrays per axis as (left, right) pairs (the
    reference passes ``mesh.<axis>.{lbc,rbc}.sigma``), ``msig`` the mesh sigma
    per axis (length = padded axis size) used by dWaveOp.  ``src`` / ``rec`` /
    ``adj`` are (idx, chan, weight) entry lists of the point operators.
    """

    def __init__(self, shape, deltas, dt, nsteps, C, kpml, msig, src, amp, rec, n_rec, adj=None,
                 order=8, pad=None):
        self.shape = tuple(int(s) for s in shape)
        self.dim = len(self.shape)
        self.deltas = tuple(float(d) for d in deltas)
        self.dt = float(dt)
        self.nsteps = int(nsteps)
        self.C = _arr(C, np.float64).reshape(-1)
        self.kpml = [(_arr(l, np.float64).reshape(-1), _arr(r, np.float64).reshape(-1)) for l, r in kpml]
        self.msig = [_arr(s, np.float64).reshape(-1) for s in msig]
        self.src = _entries(src)
        self.amp = _arr(amp, np.float64)
        self.n_src_chan = self.amp.shape[1]
        self.rec = _entries(rec)
        self.n_rec = int(n_rec)
        self.adj = _entries(adj) if adj is not None else None
        self.order = int(order)
        self.pad = pad  # ((lx, rx), ..., (lz, rz)) padding widths for the fold

    def _common(self):
        if self.dim == 2:
            nx, nz = self.shape
            ny = 1
            dx, dz = self.deltas
            dy = 1.0
            (xl, xr), (zl, zr) = self.kpml
            yl = yr = np.zeros(0)
            sx, sz = self.msig
            sy = np.zeros(1)
        else:
            nx, ny, nz = self.shape
            dx, dy, dz = self.deltas
            (xl, xr), (yl, yr), (zl, zr) = self.kpml
            sx, sy, sz = self.msig
        self._keep = (xl, xr, yl, yr, zl, zr, sx, sy, sz)
        return [self.dim, self.order, nx, ny, nz, dx, dy, dz, self.dt, self.nsteps,
                _ptr(self.C, _dp),
                _ptr(xl, _dp), xl.size, _ptr(xr, _dp), xr.size,
                _ptr(yl, _dp), yl.size, _ptr(yr, _dp), yr.size,
                _ptr(zl, _dp), zl.size, _ptr(zr, _dp), zr.size,
                _ptr(sx, _dp), _ptr(sy, _dp), _ptr(sz, _dp),
                len(self.src[0]), _ptr(self.src[0], _lp), _ptr(self.src[1], _ip), _ptr(self.src[2], _dp),
                self.n_src_chan, _ptr(self.amp, _dp)]

    @property
    def npts(self):
        return int(np.prod(self.shape))

    def forward(self, simdata=True, dwaveop=False, wavefield=False, imaging_period=1, fast=False):
        L = lib(fast)
        N = self.npts
        sd = np.zeros((self.nsteps, self.n_rec)) if simdata else None
        nimg = (self.nsteps + imaging_period - 1) // imaging_period
        dw = np.zeros((nimg, N)) if dwaveop else None
        wf = np.zeros((self.nsteps, N)) if wavefield else None
        args = self._common() + [len(self.rec[0]), _ptr(self.rec[0], _lp), _ptr(self.rec[1], _ip),
                                 _ptr(self.rec[2], _dp), self.n_rec,
                                 _ptr(sd, _dp) if sd is not None else None, imaging_period,
                                 _ptr(dw, _dp) if dw is not None else None,
                                 _ptr(wf, _dp) if wf is not None else None]
        rc = L.orc_forward(*args)
        if rc != 0:
            raise RuntimeError(f"orc_forward failed ({rc})")
        return {"simdata": sd, "dWaveOp": dw, "wavefield": wf}

    def migrate(self, operand, imaging_period=1, ckpt_interval=0, fold=True, fast=False):
        """-dt*p*sum q_k*dWaveOp_k, folded onto the unpadded grid if ``fold``."""
        if self.adj is None:
            raise ValueError("no adjoint injection entries")
        L = lib(fast)
        op = _arr(operand, np.float64)
        if op.shape != (self.nsteps, self.n_rec):
            raise ValueError(f"operand shape {op.shape} != {(self.nsteps, self.n_rec)}")
        ic = np.zeros(self.npts)
        args = self._common() + [len(self.adj[0]), _ptr(self.adj[0], _lp), _ptr(self.adj[1], _ip),
                                 _ptr(self.adj[2], _dp), self.n_rec, _ptr(op, _dp),
                                 imaging_period, ckpt_interval, _ptr(ic, _dp)]
        rc = L.orc_migrate(*args)
        if rc != 0:
            raise RuntimeError(f"orc_migrate failed ({rc})")
        if not fold:
            return ic
        return fold_padding(ic, self.shape, self.pad)


def _entries(e):
    idx, chan, w = e
    return (np.ascontiguousarray(idx, dtype=np.int64), np.ascontiguousarray(chan, dtype=np.int32),
            np.ascontiguousarray(w, dtype=np.float64))


def fold_padding(a_padded, shape, pad):
    """add_padding of model_parameter.py:537-593 on a padded flat array."""
    a = np.array(a_padded, dtype=np.float64).reshape(-1)
    dim = len(shape)
    if dim == 2:
        nx, nz = shape
        ny = 1
        (lx, rx), (lz, rz) = pad
        ly = ry = 0
    else:
        nx, ny, nz = shape
        (lx, rx), (ly, ry), (lz, rz) = pad
    out_shape = (nx - lx - rx, ny - ly - ry, nz - lz - rz)
    out = np.zeros(int(np.prod(out_shape)))
    lib().orc_fold(dim, nx, ny, nz, lx, rx, ly, ry, lz, rz, _ptr(a, _dp), _ptr(out, _dp))
    return out if dim == 3 else out


def interp_zero_fill(ts_old, data, ts_new):
    """Linear interpolation along axis 0 with zero fill outside [ts_old[0], ts_old[-1]]."""
    ts_old = np.asarray(ts_old, dtype=np.float64)
    ts_new = np.asarray(ts_new, dtype=np.float64)
    data = np.asarray(data, dtype=np.float64)
    out = np.zeros((ts_new.size,) + data.shape[1:])
    inside = (ts_new >= ts_old[0]) & (ts_new <= ts_old[-1])
    t = ts_new[inside]
    j = np.clip(np.searchsorted(ts_old, t, side="left"), 1, ts_old.size - 1)
    t0, t1 = ts_old[j - 1], ts_old[j]
    slope = (data[j] - data[j - 1]) / (t1 - t0)[:, None]
    out[inside] = slope * (t - t0)[:, None] + data[j - 1]
    return out


def misfit(residual, dt):
    return 0.5 * float(np.linalg.norm(residual) ** 2) * dt
