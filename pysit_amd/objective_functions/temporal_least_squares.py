"""Least-squares misfit in the time domain.

Mirrors pysit/objective_functions/temporal_least_squares.py:27-249:

* residual r = interp(d_obs onto solver.ts()) - S u   (:27-65)
* evaluate: J = 1/2 * dt * sum_shots ||r||^2          (:67-82)
* compute_gradient: grad = -sum_shots migrate_shot(r), with
  aux_info 'objective_value' = 1/2 dt sum ||r||^2 and 'residual_norm'
  (:141-198)
* apply_hessian ('approximate' / 'levenberg'): sum_shots F*F m1; ('full'):
  sum_shots of PySIT's two second-order migrations (:200-249, see
  ``apply_hessian``)

All shots of a call share one fused device loop per time step (the
reference loops over shots); the per-shot results are identical to running
``TemporalModeling`` shot by shot.  The batched evaluate / compute_gradient
keep the traces and residuals on the device: the observed data are uploaded
once per shot and interpolated onto each trial time axis there
(:class:`ObservedOnDevice`).  Shot-parallel MPI (the reference's
``parallel_wrap_shot``) maps to one process per GPU; with a single process
the wrap is a no-op.
"""
import numpy as np

from ..device import shots_per_batch
from ..modeling.temporal_modeling import TemporalModeling

__all__ = ["TemporalLeastSquares", "ObservedOnDevice"]


class ObservedOnDevice(object):
    """Observed traces of shots on the device, interpolated onto a time
    axis as ``ReceiverBase.interpolate_data`` does on the host
    (receivers.py:75-105): scipy's interp1d(kind='linear', bounds_error=False,
    fill_value=0) restated with the same operations in the same order
    (interp1d._call_linear: bracketing samples by searchsorted clipped to
    [1, n-1], slope (y_hi - y_lo) / (x_hi - x_lo), value slope (t - x_lo) +
    y_lo; zero outside [x_0, x_-1]).  Receivers with a time window take the
    host code.  The data are uploaded once per shot and again when they
    change through the receivers (``receivers.data = ...`` or sampling into
    them, which bump ``receivers.data_version``).  An edit of the array in
    place (``receivers.data += noise``) is not seen: call :meth:`invalidate`
    after one.  The bracketing indices are kept per time axis."""

    def __init__(self):
        self._data = {}
        self._plans = {}

    def invalidate(self, shot=None):
        """Forget the device copy of ``shot``'s observed data (all shots if
        None), so the next call uploads the current array."""
        if shot is None:
            self._data.clear()
        else:
            self._data.pop(id(shot), None)

    def __call__(self, shot, ts, dev):
        import torch
        rec = shot.receivers
        if rec._time_window[0] not in ("None", "Set"):
            # a time window multiplies the interpolated array as the host
            # code does (receivers.py:95-96): take that path itself
            return torch.as_tensor(np.asarray(rec.interpolate_data(ts), dtype=np.float64), device=dev)
        d = rec.data
        ver = getattr(rec, "data_version", 0)
        ent = self._data.get(id(shot))
        if ent is None or ent[0] is not d or ent[1] is not shot or ent[3] != ver:
            ent = (d, shot, torch.as_tensor(np.ascontiguousarray(d, dtype=np.float64), device=dev), ver)
            self._data[id(shot)] = ent
        y = ent[2]
        x = np.asarray(rec.ts, dtype=np.float64)
        ts = np.asarray(ts, dtype=np.float64)
        key = (id(shot), x.shape[0], x[0], x[-1], ts.shape[0], ts[0], ts[-1])
        plan = self._plans.get(key)
        if plan is None:
            hi = np.searchsorted(x, ts).clip(1, len(x) - 1)
            lo = hi - 1
            inside = ~((ts < x[0]) | (ts > x[-1]))
            t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a, dtype=dt), device=dev)
            plan = (t(lo, np.int64), t(hi, np.int64), t(x[lo], np.float64), t(x[hi], np.float64),
                    t(ts, np.float64), t(inside, np.bool_))
            if len(self._plans) > 4096:
                self._plans.clear()
            self._plans[key] = plan
        lo, hi, x_lo, x_hi, tt, inside = plan
        y_lo = y.index_select(0, lo)
        y_hi = y.index_select(0, hi)
        slope = (y_hi - y_lo) / (x_hi - x_lo)[:, None]
        out = slope * (tt - x_lo)[:, None] + y_lo
        return torch.where(inside[:, None], out, torch.zeros((), dtype=out.dtype, device=out.device))


class _NullWrap(object):
    use_parallel = False
    comm = None
    size = 1
    rank = 0


class TemporalLeastSquares(object):
    def __init__(self, solver, parallel_wrap_shot=None, imaging_period=1, batch_size=None,
                 memory_budget=None):
        self.solver = solver
        self.modeling_tools = TemporalModeling(solver, memory_budget=memory_budget)
        self.parallel_wrap_shot = parallel_wrap_shot or _NullWrap()
        self.imaging_period = int(imaging_period)
        self.batch_size = batch_size
        self.observed = ObservedOnDevice()

    def use_parallel(self):
        return self.parallel_wrap_shot.use_parallel

    def _groups(self, shots):
        # by default as many shots per batch as one launch can address
        step = self.batch_size or shots_per_batch(self.solver.mesh.shape(include_bc=True, as_grid=True),
                                                  np.dtype(self.solver.dtype).itemsize)
        for i in range(0, len(shots), step):
            yield shots[i:i + step]

    def _residual(self, shot, m0, dWaveOp=None, wavefield=None):
        rp = ["simdata"]
        if dWaveOp is not None:
            rp.append("dWaveOp")
        if wavefield is not None:
            rp.append("wavefield")
        rv = self.modeling_tools.forward_model(shot, m0, self.imaging_period, return_parameters=rp)
        resid = shot.receivers.interpolate_data(self.solver.ts()) - rv["simdata"]
        if dWaveOp is not None:
            dWaveOp[:] = rv["dWaveOp"][:]
        if wavefield is not None:
            wavefield[:] = rv["wavefield"][:]
        return resid

    def _allreduce(self, value):
        w = self.parallel_wrap_shot
        if not w.use_parallel:
            return value
        out = np.zeros_like(np.asarray(value, dtype=np.float64))
        w.comm.Allreduce(np.asarray(value, dtype=np.float64), out)
        return out

    def evaluate(self, shots, m0, **kwargs):
        import torch
        r2 = None
        for group in self._groups(shots):
            sims = self.modeling_tools.forward_traces_batch(group, m0, self.imaging_period)
            ts = self.solver.ts()
            for shot, sim in zip(group, sims):
                r = self.observed(shot, ts, sim.device) - sim
                s2 = torch.sum(r * r)
                r2 = s2 if r2 is None else r2 + s2
        r2 = float(self._allreduce(np.array(0.0 if r2 is None else float(r2))))
        return 0.5 * r2 * self.solver.dt

    def _gradient_helper(self, shot, m0, ignore_minus=False, **kwargs):
        dWaveOp = []
        r = self._residual(shot, m0, dWaveOp=dWaveOp)
        g = self.modeling_tools.migrate_shot(shot, m0, r, self.imaging_period, dWaveOp=dWaveOp)
        if not ignore_minus:
            g = -1 * g
        return g, r

    def compute_gradient(self, shots, m0, aux_info={}, **kwargs):
        grad = m0.perturbation()
        r2 = 0.0
        import torch
        want_h = "pseudo_hess_diag" in aux_info and aux_info["pseudo_hess_diag"][0]
        hdiag = None
        for group in self._groups(shots):
            ts_holder = {}

            def residual(i, sim, group=group):
                ts = ts_holder.setdefault("ts", self.solver.ts())
                return self.observed(group[i], ts, sim.device) - sim

            g, rs = self.modeling_tools.gradient_batch(group, m0, residual, self.imaging_period, device=True,
                                                       dw_squares=want_h)
            grad -= g
            r2 += float(sum(torch.sum(r * r) for r in rs))
            if want_h:
                # Shin (2001) illumination: sum over shots and imaged steps of
                # the physical part of dWaveOp^2 (temporal_least_squares.py:121-139)
                b = self.modeling_tools.last_batch
                pads = self.solver.mesh.pad_widths()
                for i in range(len(group)):
                    h = b.dw_squares_unpadded(i, pads)
                    hdiag = h.clone() if hdiag is None else hdiag + h
        if self.parallel_wrap_shot.use_parallel:
            r2 = float(self._allreduce(np.array(r2)))
            grad = m0.perturbation(data=self._allreduce(grad.asarray()))
        r2 *= self.solver.dt
        if "residual_norm" in aux_info and aux_info["residual_norm"][0]:
            aux_info["residual_norm"] = (True, np.sqrt(r2))
        if "objective_value" in aux_info and aux_info["objective_value"][0]:
            aux_info["objective_value"] = (True, 0.5 * r2)
        if want_h:
            h = hdiag.cpu().numpy().reshape(-1, 1) * self.imaging_period
            if self.parallel_wrap_shot.use_parallel:
                h = self._allreduce(h)
            aux_info["pseudo_hess_diag"] = (True, h * self.solver.dt)
        return grad

    def apply_hessian(self, shots, m0, m1, hessian_mode="approximate", levenberg_mu=0.0, *args, **kwargs):
        modes = ["approximate", "full", "levenberg"]
        if hessian_mode not in modes:
            raise ValueError(f"Invalid Hessian mode.  Valid options for applying hessian are {modes}")
        tm = self.modeling_tools
        result = m0.perturbation()
        if hessian_mode == "full":
            for shot in shots:
                result += self._full_hessian_shot(shot, m0, m1)
            if self.parallel_wrap_shot.use_parallel:
                result = m0.perturbation(data=self._allreduce(result.asarray()))
            return result
        for shot in shots:
            dW0 = tm.forward_model(shot, m0, 1, return_parameters=["dWaveOp"])["dWaveOp"]
            d1 = tm.linear_forward_model(shot, m0, m1, return_parameters=["simdata"], dWaveOp0=dW0)["simdata"]
            result += tm.migrate_shot(shot, m0, d1, 1, dWaveOp=dW0)
        if self.parallel_wrap_shot.use_parallel:
            result = m0.perturbation(data=self._allreduce(result.asarray()))
        if hessian_mode == "levenberg":
            result += levenberg_mu * m1
        return result

    def full_hessian_terms(self, shot, m0, m1):
        """The two migrations of PySIT's full Hessian for one shot
        (temporal_least_squares.py:224-239), every step imaged:

        * res1 = migrate_shot(r0, dWaveOp=dWaveOp1): the adjoint q of the
          residual r0 imaged against the dWaveOp of the Born field u1 of m1,
          keeping dWaveOpAdj1, the dWaveOp of q;
        * res2 = migrate_shot(d1, operand_dWaveOpAdj=dWaveOpAdj1,
          operand_model=m1, dWaveOp=dWaveOp0): the adjoint driven by the Born
          traces d1 and the volume source m1 * dWaveOpAdj1_k.

        Returns (res1, res2).  PySIT's 'full' mode is res1 + res2.  The
        derivative of the gradient -migrate_shot(r) along m1 is res2 - res1:
        d(dWaveOp0) = dWaveOp1 enters with the gradient's sign, and res1 is
        built from +r0."""
        tm = self.modeling_tools
        # imaging period 1 throughout: dWaveOp0[k], dWaveOp1[k] and
        # dWaveOpAdj1[k] are indexed by the time step k
        rv = tm.forward_model(shot, m0, 1, return_parameters=["simdata", "dWaveOp"])
        dW0 = rv["dWaveOp"]
        r0 = shot.receivers.interpolate_data(self.solver.ts()) - rv["simdata"]
        lin = tm.linear_forward_model(shot, m0, m1, return_parameters=["simdata", "dWaveOp1"], dWaveOp0=dW0)
        adj1 = []
        res1 = tm.migrate_shot(shot, m0, r0, 1, dWaveOp=lin["dWaveOp1"], dWaveOpAdj=adj1)
        res2 = tm.migrate_shot(shot, m0, lin["simdata"], 1, operand_dWaveOpAdj=adj1, operand_model=m1, dWaveOp=dW0)
        return res1, res2

    def _full_hessian_shot(self, shot, m0, m1):
        res1, res2 = self.full_hessian_terms(shot, m0, m1)
        return res1 + res2
