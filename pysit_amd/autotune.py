"""Launch-configuration autotuner for device batches.

The step kernels of a batch run in several configurations that give the
same results bit for bit and differ only in speed, which depends on the
grids and the number of shots:

* ``tile_x``: rows along x per wavefront tile (2D: 8, 4 or 2; short tiles give
  more wavefronts and shorter marches, which suits small batches);
* ``tile_order``: the tile table order, ``'shot'`` or, for shots of one
  shape and length, ``'interleave'``;
* ``graph_steps``: the time loop as replayed hipGraphs of GRAPH_STEPS (64) steps, or one
  launch per step.

:func:`tuned_batch` builds the batch in each candidate configuration, times a
short forward-with-dWaveOp plus adjoint-imaging pass on the device
(:meth:`DeviceBatch.probe_step_time`) and returns the batch in the fastest
one; a candidate must beat the default (or the best so far) by ``MARGIN`` to
be chosen, so near ties resolve the same way in every run.  Choices are cached per signature -- dimension, dtype, order, imaging
period, FP32 window and every shot's padded grid, PML widths and receiver
count, not the step count, so an inversion whose dt changes keeps its choice
-- in memory and, when the environment variable ``PYSIT_AMD_TUNE_CACHE``
names a JSON file, in that file.
"""
import json
import os

import numpy as np

from . import _native as N
from .device import GRAPH_STEPS, DeviceBatch

__all__ = ["tuned_batch", "signature", "candidates", "clear_cache"]

_CACHE = {}
# a candidate replaces the one before it (the default first) only when it is
# this much faster: near ties keep the default, so repeated runs choose alike
MARGIN = 0.03
# device memory a probe's dWaveOp store may take (its slots are cut to fit,
# down to 16); fewer slots than two graph chunks time single launches only
PROBE_BYTES = 4 << 30


def signature(shots, dtype, order, imaging_period=1, ic_window=None):
    """Cache key of a batch (a string)."""
    parts = [len(shots[0].shape), np.dtype(dtype).name, int(order), int(imaging_period),
             -1 if ic_window is None else int(ic_window)]
    for g in shots:
        pml = [[len(np.atleast_1d(a)) for a in pair] for pair in g.kpml]
        parts.append([list(g.shape), pml, int(g.nrec)])
    return json.dumps(parts, separators=(",", ":"))


def candidates(shots, dtype):
    """Tile configurations worth timing for these shots (graph steps are
    tried afterwards, for the fastest)."""
    dim = len(shots[0].shape)
    code = 1 if np.dtype(dtype) == np.float64 else 0
    rows = [t for t in (N.tile_rows(dim, code, 0), 4, 2, 8) if N.tile_rows_supported(dim, code, t)]
    rows = list(dict.fromkeys(rows))
    alike = len({tuple(g.shape) for g in shots}) == 1 and len({g.nt for g in shots}) == 1
    # alike shots: the interleaved order first (the shots of a tile share
    # its model rows in L2; measured 2-3 % ahead), so it is the default that
    # the others must beat
    orders = ["interleave", "shot"] if alike and len(shots) > 1 else ["shot"]
    return [{"tile_x": t, "tile_order": o} for t in rows for o in orders]


def clear_cache():
    _CACHE.clear()


def _file():
    return os.environ.get("PYSIT_AMD_TUNE_CACHE")


def _load(sig):
    if sig in _CACHE:
        return _CACHE[sig]
    f = _file()
    if f and os.path.exists(f):
        try:
            with open(f) as fh:
                cfg = json.load(fh).get(sig)
        except (OSError, ValueError):
            cfg = None
        if cfg is not None:
            _CACHE[sig] = cfg
            return cfg
    return None


def _store(sig, cfg):
    _CACHE[sig] = cfg
    f = _file()
    if not f:
        return
    try:
        data = {}
        if os.path.exists(f):
            with open(f) as fh:
                data = json.load(fh)
        data[sig] = cfg
        tmp = f + ".tmp"
        with open(tmp, "w") as fh:
            json.dump(data, fh)
        os.replace(tmp, f)
    except (OSError, ValueError):
        pass


def _faster(t, best_t):
    return t < best_t * (1.0 - MARGIN)


def tuned_batch(shots, model, deltas, dt, order, dtype=np.float32, slots=2 * GRAPH_STEPS, report=None, **kw):
    """A :class:`DeviceBatch` of ``shots`` in the fastest launch
    configuration (keyword arguments other than the tuned ones pass through).
    ``report``, a list, receives (configuration, seconds per slot) of every
    candidate timed.  Every tile candidate is timed over the same ``slots``
    slots (fewer where their dWaveOp store would pass PROBE_BYTES) with
    single launches and with graphs: at least two graph chunks, so the graph
    pass replays its graphs (its first pass captures them).
    Small batches are bound by launch cost without graphs, which hides the
    tile heights' difference; with graphs it shows."""
    for k in ("tile_x", "tile_order", "graph_steps"):
        kw.pop(k, None)
    sig = signature(shots, dtype, order, kw.get("imaging_period", 1), kw.get("ic_window"))
    cfg = _load(sig)
    if cfg is None:
        import torch
        timed = []
        for cand in candidates(shots, dtype):
            b = DeviceBatch(shots, model, deltas, dt, order, dtype=dtype, graph_steps=0, **cand, **kw)
            S = min(int(slots), max(16, PROBE_BYTES // (b.total_n * b.dtype.itemsize)))
            for gs in ((0, GRAPH_STEPS) if S >= 2 * GRAPH_STEPS else (0,)):
                b.graph_steps = gs
                timed.append((b.probe_step_time(S), dict(cand, graph_steps=gs)))
            del b
        best_t, best = timed[0]
        for t, c in timed[1:]:
            if _faster(t, best_t):
                best_t, best = t, c
        torch.cuda.empty_cache()
        cfg = best
        if report is not None:
            report.extend((c, t) for t, c in timed)
        _store(sig, cfg)
    return DeviceBatch(shots, model, deltas, dt, order, dtype=dtype, **cfg, **kw)
