"""Time-domain modeling on the GPU with PySIT's ``TemporalModeling`` API.

Mirrors pysit/modeling/temporal_modeling.py: ``forward_model`` (:60-169),
``migrate_shot`` (:171-214), ``adjoint_model`` (:228-387) and
``linear_forward_model`` (:389-546), with the same arguments and return
dictionaries.  The per-step Python loops of the reference become one fused
device loop per call (``pysit_amd.device.DeviceBatch``):

* forward: traces are sampled from u_k before each step, rhs_k = S* w(k dt),
  dWaveOp_k (the imaging factor) is produced by the step kernel itself;
* adjoint: q runs the same step backwards in time driven by S* r_k, the FP64
  imaging accumulator gathers q_k * dWaveOp_k for k % imaging_period == 0,
  and the result is folded onto the physical grid with the factor
  -dt * imaging_period;
* Born: u1 is driven by -m1 * dWaveOp0_k.

``forward_model(..., return_parameters=['dWaveOp'])`` returns a list of
lightweight device handles instead of host arrays (one per imaged step).
``TemporalLeastSquares`` copies that list and hands it back to
``migrate_shot``, which then uses the device history (or its checkpoints)
directly.  Indexing a handle with ``numpy.asarray`` downloads that step.

Batched extensions run many shots in one launch per time step:
``forward_model_batch`` and ``gradient_batch``.
"""
import numpy as np

from ..core.receivers import adjoint_entries, sampling_entries
from ..device import DeviceBatch, ShotGeometry

__all__ = ["TemporalModeling", "DWaveOpHandle"]


class DWaveOpHandle(object):
    """dWaveOp of one imaged step of one shot, held on the device: of the
    forward field (``store`` "dw"), of the Born field ("dw1") or of the
    adjoint field ("dw_adj")."""

    __slots__ = ("batch", "shot", "k", "store")

    def __init__(self, batch, shot, k, store="dw"):
        self.batch = batch
        self.shot = shot
        self.k = k
        self.store = store

    def __array__(self, dtype=None, copy=None):
        a = self.batch.shot_dw(self.shot, self.k, self.store).reshape(-1, 1).double().cpu().numpy()
        return a if dtype is None else a.astype(dtype)


def _handles(batch, shot, nsteps, p, store="dw"):
    return [DWaveOpHandle(batch, shot, k, store) for k in range(0, nsteps, p)]


def _on_device(seq, batch=None, shot=None):
    """The handles' (batch, shot, store) when ``seq`` is a list of device
    handles (of ``batch``/``shot`` when given), else None."""
    if seq is None or not len(seq) or not isinstance(seq[0], DWaveOpHandle):
        return None
    h = seq[0]
    if (batch is not None and h.batch is not batch) or (shot is not None and h.shot != shot):
        return None
    return h.batch, h.shot, h.store


class TemporalModeling(object):
    """Forward, adjoint, migration and Born modeling for a time solver."""

    solver_type = "time"
    modeling_type = "time"

    def __init__(self, solver, memory_budget=None, tile_x=None, ic_window=None, autotune=True, adjoint="pysit"):
        if self.solver_type != solver.supports["equation_dynamics"]:
            raise TypeError(f"Argument 'solver' type {solver.supports['equation_dynamics']} does not match "
                            f"modeling solver type {self.solver_type}.")
        self.solver = solver
        self.memory_budget = memory_budget
        self.tile_x = tile_x
        self.ic_window = ic_window  # None: the batch default (FP32: IC_WINDOW = 64 slots)
        # launch configuration (tile rows, tile order, graphs) picked by timing
        # the candidates once per batch signature, unless tile_x is given
        self.autotune = bool(autotune)
        # "pysit": dWaveOp as the reference forms it (mesh PML profile), so
        # migrate_shot and the Born operator are PySIT's; "exact": the step's
        # own m-derivative (kernel PML profile), so migrate_shot is the exact
        # adjoint of the discrete forward map and linear_forward_model its
        # exact linearisation (2D; see DeviceBatch)
        if adjoint not in ("pysit", "exact"):
            raise ValueError("adjoint must be 'pysit' or 'exact'")
        self.adjoint = adjoint

    # ---------------------------------------------------------------- plumbing
    def geometry(self, shot):
        """The device description of ``shot`` for the current model.  The
        receiver entries do not depend on the time axis: they are kept per
        sampling operator and grid, so the copies of one receiver set (a
        survey's shots share its operator) gather them once."""
        s = self.solver
        shape = tuple(s.mesh.shape(include_bc=True, as_grid=True))
        src, amp = shot.sources.injection(s.dt, s.nsteps)
        cache = self.__dict__.setdefault("_rec_entries", {})
        op = shot.receivers.sampling_operator
        op = shot.receivers if op is None else op
        key = (id(op), shape)
        hit = cache.get(key)
        if hit is None or hit[0] is not op:
            if len(cache) > 4096:
                cache.clear()
            hit = (op, sampling_entries(shot.receivers, shape), adjoint_entries(shot.receivers))
            cache[key] = hit
        rec, adj = hit[1], hit[2]
        return ShotGeometry(shape, s.kernel_pml(), s.model_window(), s.nsteps, src, amp, rec,
                            shot.receivers.receiver_count, adj=adj)

    def batch(self, shots, imaging_period=1, reuse=False):
        """The device batch of ``shots`` for the solver's current model.  With
        ``reuse`` the last reusable batch is kept: the same shots on the same
        time axis (the evaluations of an optimisation loop) reuse its tables,
        state and stores and only take the new model.  Per-shot calls never
        reuse, so device handles they return keep their own model and
        history."""
        s = self.solver
        key = (tuple(id(sh) for sh in shots), s.dt, s.nsteps, imaging_period, np.dtype(s.dtype).str,
               s.spatial_accuracy_order, tuple(s.mesh.shape(include_bc=True, as_grid=True)), self.adjoint)
        cached = getattr(self, "_batch_cache", None)
        if reuse and cached is not None and cached[0] == key:
            b = cached[1]
            b.set_model(s.unpadded_model())
            return b
        if reuse and cached is not None and cached[0][:1] + cached[0][3:] == key[:1] + key[3:]:
            # the same shots on another time axis (a trial model with another
            # largest speed): new amplitudes and step counts, same tables
            b = cached[1]
            try:
                b.retime([self.geometry(sh) for sh in shots], s.dt)
            except ValueError:
                b = None
            if b is not None:
                b.set_model(s.unpadded_model())
                self._batch_cache = (key, b, list(shots))
                return b
        if reuse:
            self._batch_cache = None
        b = self.make_batch([self.geometry(sh) for sh in shots], s.unpadded_model(), imaging_period)
        if reuse:
            self._batch_cache = (key, b, list(shots))
        return b

    def make_batch(self, geometries, model, imaging_period=1):
        """A DeviceBatch of ``geometries`` for ``model`` (unpadded wave speed,
        host array or device tensor) on the solver's time axis, autotuned
        unless disabled or tile_x is fixed."""
        s = self.solver
        kw = dict(dtype=s.dtype, imaging_period=imaging_period, memory_budget=self.memory_budget,
                  ic_window=self.ic_window, adjoint=self.adjoint)
        if self.autotune and not self.tile_x:
            from ..autotune import tuned_batch
            return tuned_batch(geometries, model, s.mesh.deltas, s.dt, s.spatial_accuracy_order, **kw)
        return DeviceBatch(geometries, model, s.mesh.deltas, s.dt, s.spatial_accuracy_order, tile_x=self.tile_x,
                           **kw)

    def _unpad_field(self, padded):
        mesh = self.solver.mesh
        return mesh.unpad_array(np.asarray(padded, dtype=np.float64).reshape(-1, 1), copy=True)

    def _fold_images(self, batch, scale, shots_idx=None):
        """Sum over shots of fold(scale * image) as an unpadded Perturbation."""
        import torch
        s = self.solver
        mesh = s.mesh
        out = torch.zeros(tuple(mesh.shape(as_grid=True)), dtype=torch.float64, device=batch.dev)
        if shots_idx is None:
            batch.fold_sum(mesh.pad_widths(), out, scale=scale)
        else:
            for i in shots_idx:
                batch.fold(i, mesh.pad_widths(), out, scale=scale, accumulate=True)
        return s.model_parameters.without_padding().perturbation(data=out.cpu().numpy().reshape(-1, 1))

    # ----------------------------------------------------------------- forward
    def forward_model(self, shot, m0, imaging_period=1, return_parameters=[]):
        solver = self.solver
        solver.model_parameters = m0
        keep = "dWaveOp" in return_parameters
        b = self.batch([shot], imaging_period)
        traces = b.forward(keep_dw=keep, store_field="wavefield" in return_parameters)
        retval = dict()
        if "wavefield" in return_parameters:
            retval["wavefield"] = [self._unpad_field(b.shot_field(b.field_store, 0, k).cpu().numpy())
                                   for k in range(solver.nsteps)]
            b.field_store = None
        if keep:
            retval["dWaveOp"] = _handles(b, 0, solver.nsteps, imaging_period)
        if "simdata" in return_parameters:
            retval["simdata"] = traces[0].cpu().numpy()
        return retval

    def forward_model_batch(self, shots, m0, imaging_period=1, keep_dw=False):
        """Traces of many shots in one fused loop (host arrays, shot order).
        With ``keep_dw`` the batch is kept in ``self.last_batch`` for
        :meth:`gradient_batch`-style reuse."""
        self.solver.model_parameters = m0
        b = self.batch(shots, imaging_period, reuse=True)
        traces = b.forward(keep_dw=keep_dw)
        self.last_batch = b if keep_dw else None
        return [t.cpu().numpy() for t in traces]

    # ----------------------------------------------------------------- adjoint
    def _batch_for(self, shot, dWaveOp, imaging_period):
        """Reuse the device history behind handles, or upload host arrays.
        Returns (batch, shot index in it, name of the history)."""
        s = self.solver
        dev = _on_device(dWaveOp)
        if dev is not None:
            b, si, store = dev
            if (b.dt == s.dt and b.max_nt == s.nsteps and b.imaging_period == imaging_period
                    and b.adjoint == self.adjoint and len(dWaveOp) == len(range(0, s.nsteps, imaging_period))):
                return b, si, store
            raise ValueError("dWaveOp handles belong to another model, time axis, imaging period or adjoint mode")
        b = self.batch([shot], imaging_period)
        if dWaveOp is not None:
            b.load_dw([list(dWaveOp)])
        return b, 0, "dw"

    def _history(self, b, si, seq):
        """A (steps, total_n) device history of ``b`` holding ``seq`` (one
        padded array or handle per step) for shot ``si``: the store itself
        when ``seq`` are handles of it, else an upload."""
        dev = _on_device(seq, b, si)
        if dev is not None:
            hist = getattr(b, dev[2], None)
            if hist is None:
                raise RuntimeError(f"the {dev[2]} history behind these handles is no longer held")
            return hist
        return b.history_from(si, [np.asarray(a) for a in seq])

    def adjoint_model(self, shot, m0, operand_simdata, imaging_period=1, operand_dWaveOpAdj=None,
                      operand_model=None, return_parameters=[], dWaveOp=None, wavefield=None):
        solver = self.solver
        solver.model_parameters = m0
        do_ic = dWaveOp is not None
        if not do_ic and "imaging_condition" in return_parameters:
            raise ValueError("To compute imaging condition, forward component must be specified.")
        want_adj = "dWaveOpAdj" in return_parameters
        # the reference adds the volume source only when both are given
        # (_setup_adjoint_rhs, temporal_modeling.py:221-226)
        vol = operand_dWaveOpAdj is not None and operand_model is not None
        if (want_adj or vol) and imaging_period != 1:
            raise ValueError("dWaveOpAdj and operand_dWaveOpAdj on the device need imaging_period 1")
        b, si, image_store = self._batch_for(shot, dWaveOp, imaging_period)
        operands = [None] * b.n_shots
        for i in range(b.n_shots):
            operands[i] = np.zeros((b.shots[b.order_idx.index(i)].nt, b.shots[b.order_idx.index(i)].nrec))
        operands[si] = np.asarray(operand_simdata, dtype=np.float64)
        volume = None
        if vol:
            # operand_model.with_padding(): zero padding (mesh.pad_array's default)
            m1p = np.asarray(operand_model.with_padding().data)
            volume = (b.padded_operand(si, m1p), self._history(b, si, operand_dWaveOpAdj))
        store = "adjointfield" in return_parameters
        b.migrate(operands, store_field=store, image=do_ic, image_store=image_store, store_adjoint_dw=want_adj,
                  volume=volume)
        retval = dict()
        if store:
            retval["adjointfield"] = [self._unpad_field(b.shot_field(b.adjoint_store, si, k).cpu().numpy())
                                      for k in range(solver.nsteps)]
            b.adjoint_store = None
        if want_adj:
            retval["dWaveOpAdj"] = _handles(b, si, solver.nsteps, 1, "dw_adj")
        if do_ic:
            retval["imaging_condition"] = self._fold_images(b, -solver.dt * imaging_period, [si])
        return retval

    def migrate_shot(self, shot, m0, operand_simdata, imaging_period, operand_dWaveOpAdj=None,
                     operand_model=None, dWaveOp=None, adjointfield=None, dWaveOpAdj=None,
                     wavefield=None):
        if dWaveOp is None:
            dWaveOp = self.forward_model(shot, m0, imaging_period, return_parameters=["dWaveOp"])["dWaveOp"]
        rp = ["imaging_condition"]
        if adjointfield is not None:
            rp.append("adjointfield")
        if dWaveOpAdj is not None:
            rp.append("dWaveOpAdj")
        rv = self.adjoint_model(shot, m0, operand_simdata, imaging_period, operand_dWaveOpAdj, operand_model,
                                return_parameters=rp, dWaveOp=dWaveOp, wavefield=wavefield)
        if adjointfield is not None:
            adjointfield[:] = rv["adjointfield"][:]
        if dWaveOpAdj is not None:
            dWaveOpAdj[:] = rv["dWaveOpAdj"][:]
        return rv["imaging_condition"].without_padding()

    def forward_traces_batch(self, shots, m0, imaging_period=1):
        """Traces of many shots in one fused loop, as device FP64 tensors in
        the shots' order (forward_model_batch without the host copies)."""
        self.solver.model_parameters = m0
        b = self.batch(shots, imaging_period, reuse=True)
        traces = b.forward(keep_dw=False)
        self.last_batch = None
        return traces

    def gradient_batch(self, shots, m0, operands_fn, imaging_period=1, device=False, dw_squares=False):
        """Forward all shots, form each shot's operand with ``operands_fn(i,
        simdata)``, migrate them together and return (sum of migrate_shot
        results, list of operands).  With ``device`` the simdata handed to
        ``operands_fn`` and the operands stay device tensors.  With
        ``dw_squares`` the batch also sums every shot's dWaveOp^2 over the
        imaged steps (``last_batch.dw_squares_unpadded``)."""
        solver = self.solver
        solver.model_parameters = m0
        b = self.batch(shots, imaging_period, reuse=True)
        traces = b.forward(keep_dw=True)
        if device:
            ops = [operands_fn(i, t) for i, t in enumerate(traces)]
        else:
            ops = [np.asarray(operands_fn(i, t.cpu().numpy()), dtype=np.float64) for i, t in enumerate(traces)]
        b.migrate(ops, dw_squares=dw_squares)
        g = self._fold_images(b, -solver.dt * imaging_period)
        self.last_batch = b
        return g, ops

    # -------------------------------------------------------------------- Born
    def linear_forward_model(self, shot, m0, m1, return_parameters=[], dWaveOp0=None):
        solver = self.solver
        solver.model_parameters = m0
        if dWaveOp0 is not None:
            b, si, store = self._batch_for(shot, dWaveOp0, 1)
            if store != "dw":
                raise ValueError("dWaveOp0 must be the forward field's dWaveOp")
        else:
            b = self.batch([shot], 1)
            b.forward(keep_dw=True, segment=solver.nsteps)
            si = 0
        m1u = m1.without_padding().data.reshape(tuple(solver.mesh.shape(as_grid=True)))
        keep1 = "dWaveOp1" in return_parameters
        want_u1 = "wavefield1" in return_parameters
        traces = b.born(m1u, store_dw=keep1, store_field=want_u1)
        retval = dict()
        if want_u1:
            retval["wavefield1"] = [self._unpad_field(b.shot_field(b.field_store, si, k).cpu().numpy())
                                    for k in range(solver.nsteps)]
            b.field_store = None
        if "simdata" in return_parameters:
            retval["simdata"] = traces[si].cpu().numpy()
        if "dWaveOp0" in return_parameters:
            retval["dWaveOp0"] = _handles(b, si, solver.nsteps, 1)
        if keep1:
            retval["dWaveOp1"] = _handles(b, si, solver.nsteps, 1, "dw1")
        return retval
