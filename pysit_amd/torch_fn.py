"""``torch.autograd`` wrapper: traces of many shots as a function of the model.

``ShotTraces.apply(m, problem)`` returns the recorded traces
``(n_shots, nt, nrec)`` (FP64) of every shot in ``problem`` for the slowness
squared ``m = C^-2`` on the unpadded model grid.  Its backward is the HIP
adjoint: the upstream trace gradient G is injected at the receivers
(S*, no prod(deltas)), propagated backwards in time, and the zero-lag imaging
condition against the stored (or checkpoint-recomputed) dWaveOp gives

    dL/dm = -prod(deltas) * fold( sum_k q_k[G] * dWaveOp_k ),

with PySIT's adjoint recursion (pysit/modeling/temporal_modeling.py:228-387).
``ShotProblem(adjoint=...)`` picks the dWaveOp it images against:

* ``"pysit"`` (default): PySIT's, built with the mesh PML profile
  (constant_density_acoustic_base.py:24-40).  The backward then equals
  PySIT's migrate_shot up to the factor -prod(deltas)/dt
  (``TemporalLeastSquares`` gradients are reproduced by
  ``L = 0.5 * dt * ||d - d_obs||^2``); it is the exact gradient without PML
  and departs from it on the right-hand PML sides, where the reference's
  kernel profile sits one node off its mesh profile.
* ``"exact"`` (2D): the step's own derivative with respect to m (kernel
  profile), which makes the backward the exact gradient of the discrete
  forward map with the PML on as well: the adjoint recursion is the step's
  transpose after the change of variables Phi' = -dt (sz - sx) mu of the
  auxiliary multipliers.

dt is fixed by the model given to the forward (CFL), as in PySIT.  A batch
serves one pending backward: a second forward on the same problem before
the first output's backward makes that backward raise.
"""
import numpy as np
import torch

from .device import DeviceBatch

__all__ = ["ShotProblem", "ShotTraces", "traces"]


class ShotProblem(object):
    """Shots of one solver / mesh for the autograd path (geometry built once
    per model change by ``TemporalModeling.geometry``)."""

    def __init__(self, solver, shots, imaging_period=1, memory_budget=None, tile_x=None, segment=None,
                 ic_window=None, autotune=True, adjoint="pysit"):
        from .modeling.temporal_modeling import TemporalModeling
        self.solver = solver
        self.shots = shots
        self.tm = TemporalModeling(solver, memory_budget=memory_budget, tile_x=tile_x, ic_window=ic_window,
                                   autotune=autotune, adjoint=adjoint)
        self.imaging_period = imaging_period
        self.segment = segment
        self.batch = None
        self._key = None
        self._m_host = None
        self.timer = None

    def prepare(self, m_host):
        s = self.solver
        C = np.sqrt(1.0 / np.asarray(m_host, dtype=np.float64)).reshape(-1, 1)
        s.model_parameters = s.ModelParameters(s.mesh, {"C": C})
        return s

    def device_batch(self, m):
        """The batch for slowness squared ``m`` (device tensor): rebuilt when
        dt or the step count change, else the model is updated in place.
        The solver takes the model only when it differs from the last one
        this problem gave it and nothing else set a model since (the host
        model rebuild is the costly part of a repeated call)."""
        m_host = m.detach().cpu().numpy()
        s = self.solver
        if not (self._m_host is not None and self._m_host[1] == s._model_change_count
                and np.array_equal(m_host, self._m_host[0])):
            s = self.prepare(m_host)
            self._m_host = (m_host.copy(), s._model_change_count)
        key = (s.dt, s.nsteps)
        C = (m.detach() ** -0.5).to(torch.float64)
        if self.batch is None or self._key != key:
            self.batch = None
            self.batch = self.tm.make_batch([self.tm.geometry(sh) for sh in self.shots], C, self.imaging_period)
            self._key = key
        else:
            self.batch.set_model(C)
        self.batch.timer = self.timer
        return self.batch


class ShotTraces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, m, problem):
        b = problem.device_batch(m)
        tr = b.forward(keep_dw=True, segment=problem.segment)
        ctx.problem = problem
        ctx.batch = b
        ctx.generation = b.generation
        ctx.m_shape = m.shape
        return torch.stack(tr)

    @staticmethod
    def backward(ctx, grad_traces):
        b = ctx.batch
        p = ctx.problem
        if b.generation != ctx.generation:
            # the batch's model, dWaveOp store and checkpoints now belong to a
            # later forward: this backward would image against that one
            raise RuntimeError("ShotTraces: the shot batch ran another forward since this output was "
                               "computed; call backward before the next forward on the same ShotProblem, "
                               "or use one ShotProblem per pending graph")
        mesh = p.solver.mesh
        b.migrate([grad_traces[i].contiguous() for i in range(b.n_shots)])
        out = torch.zeros(tuple(mesh.shape(as_grid=True)), dtype=torch.float64, device=b.dev)
        scale = -float(np.prod(mesh.deltas)) * p.imaging_period
        b.fold_sum(mesh.pad_widths(), out, scale=scale)
        return out.reshape(ctx.m_shape).to(grad_traces.dtype), None


def traces(m, problem):
    """Recorded traces (n_shots, nt, nrec) of ``problem`` for slowness squared ``m``."""
    return ShotTraces.apply(m, problem)
