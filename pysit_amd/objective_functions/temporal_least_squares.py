"""Least-squares misfit in the time domain.

Mirrors pysit/objective_functions/temporal_least_squares.py:27-249:

* residual r = interp(d_obs onto solver.ts()) - S u   (:27-65)
* evaluate: J = 1/2 * dt * sum_shots ||r||^2          (:67-82)
* compute_gradient: grad = -sum_shots migrate_shot(r), with
  aux_info 'objective_value' = 1/2 dt sum ||r||^2 and 'residual_norm'
  (:141-198)
* apply_hessian ('approximate' / 'levenberg'): sum_shots F*F m1 (:200-249)

All shots of a call share one fused device loop per time step (the
reference loops over shots); the per-shot results are identical to running
``TemporalModeling`` shot by shot.  Shot-parallel MPI (the reference's
``parallel_wrap_shot``) maps to one process per GPU; with a single process
the wrap is a no-op.
"""
import numpy as np

from ..modeling.temporal_modeling import TemporalModeling

__all__ = ["TemporalLeastSquares"]


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

    def use_parallel(self):
        return self.parallel_wrap_shot.use_parallel

    def _groups(self, shots):
        step = self.batch_size or len(shots)
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
        r2 = 0.0
        ts = None
        for group in self._groups(shots):
            sims = self.modeling_tools.forward_model_batch(group, m0, self.imaging_period)
            ts = self.solver.ts()
            for shot, sim in zip(group, sims):
                r = shot.receivers.interpolate_data(ts) - sim
                r2 += np.linalg.norm(r) ** 2
        r2 = float(self._allreduce(np.array(r2)))
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
        for group in self._groups(shots):
            ts_holder = {}

            def residual(i, sim, group=group):
                ts = ts_holder.setdefault("ts", self.solver.ts())
                return group[i].receivers.interpolate_data(ts) - sim

            g, rs = self.modeling_tools.gradient_batch(group, m0, residual, self.imaging_period)
            grad -= g
            r2 += sum(float(np.linalg.norm(r) ** 2) for r in rs)
        if self.parallel_wrap_shot.use_parallel:
            r2 = float(self._allreduce(np.array(r2)))
            grad = m0.perturbation(data=self._allreduce(grad.asarray()))
        r2 *= self.solver.dt
        if "residual_norm" in aux_info and aux_info["residual_norm"][0]:
            aux_info["residual_norm"] = (True, np.sqrt(r2))
        if "objective_value" in aux_info and aux_info["objective_value"][0]:
            aux_info["objective_value"] = (True, 0.5 * r2)
        if "pseudo_hess_diag" in aux_info and aux_info["pseudo_hess_diag"][0]:
            raise NotImplementedError("the pseudo-Hessian diagonal is not computed by the device path")
        return grad

    def apply_hessian(self, shots, m0, m1, hessian_mode="approximate", levenberg_mu=0.0, *args, **kwargs):
        modes = ["approximate", "full", "levenberg"]
        if hessian_mode not in modes:
            raise ValueError(f"Invalid Hessian mode.  Valid options for applying hessian are {modes}")
        if hessian_mode == "full":
            raise NotImplementedError("the full Hessian needs dWaveOp1/dWaveOpAdj, not provided on the device")
        tm = self.modeling_tools
        result = m0.perturbation()
        for shot in shots:
            dW0 = tm.forward_model(shot, m0, 1, return_parameters=["dWaveOp"])["dWaveOp"]
            d1 = tm.linear_forward_model(shot, m0, m1, return_parameters=["simdata"], dWaveOp0=dW0)["simdata"]
            result += tm.migrate_shot(shot, m0, d1, 1, dWaveOp=dW0)
        if self.parallel_wrap_shot.use_parallel:
            result = m0.perturbation(data=self._allreduce(result.asarray()))
        if hessian_mode == "levenberg":
            result += levenberg_mu * m1
        return result
