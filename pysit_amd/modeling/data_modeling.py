"""Observed-data generation for time solvers.

Mirrors the time path of pysit/modeling/data_modeling.py
(``generate_seismic_data`` :17-178 and ``generate_shot_data_time``
:335-429): the solver takes the model, every shot's receivers get a fresh
time axis (``reset_time_series(solver.ts())``) and the traces recorded
before each step, and ``shot.dt`` / ``shot.trange`` are set.  All shots run
in one fused device loop.

Shot data on disk (:62-131 save, ``generate_seismic_data_from_file``
:180-330 load): one file per shot for the traces and one for the time axis,
``shot_<i>`` and ``ts_<i>`` (i from 1) in ``./shots_time``:

* ``'savemat'``: MATLAB files whose variable is named by the file's own path,
  as the reference writes them (files from either side read on the other);
* ``'pickle'``: the reference's pickled arrays (``.save``), read back with an
  unpickler that only rebuilds numpy arrays;
* ``'npy'``: numpy files without pickling;
* ``'h5py'``: needs h5py, which this environment lacks (ImportError, as the
  reference raises).
"""
import os
import pickle
import re
import time

import numpy as np

from ..device import shots_per_batch
from .temporal_modeling import TemporalModeling

__all__ = ["generate_seismic_data", "generate_shot_data_time", "generate_seismic_data_from_file",
           "save_shot_data", "load_shot_data"]

_EXT = {"savemat": ".mat", "pickle": ".save", "npy": ".npy", "h5py": ".hdf5"}


class _ArrayUnpickler(pickle.Unpickler):
    """Rebuilds numpy arrays and nothing else (any other global refused).

    The globals are those numpy's own array pickles name: protocol 2
    (``_reconstruct``, ``ndarray``, ``dtype``, with ``_codecs.encode`` for the
    raw bytes) and protocol 5 (``_frombuffer``), under both numpy module
    paths.  Python 2 pickles load with ``encoding='latin1'``, as numpy
    requires for them."""
    _ALLOWED = {(m, n) for m in ("numpy.core.multiarray", "numpy._core.multiarray")
                for n in ("_reconstruct", "scalar")}
    _ALLOWED |= {("numpy.core.numeric", "_frombuffer"), ("numpy._core.numeric", "_frombuffer"),
                 ("numpy", "ndarray"), ("numpy", "dtype"), ("_codecs", "encode")}

    def __init__(self, fh):
        super().__init__(fh, encoding="latin1")

    def find_class(self, module, name):
        if (module, name) in self._ALLOWED:
            return super().find_class(module, name)
        raise pickle.UnpicklingError(f"refusing to load {module}.{name} from a shot data file")


def _paths(path, i, method):
    # the reference joins with '/' and names each MATLAB variable by this string
    ext = _EXT[method]
    return f"{path}/shot_{i + 1}{ext}", f"{path}/ts_{i + 1}{ext}"


def save_shot_data(shots, save_method, path="./shots_time"):
    """Write every shot's traces and time axis.  The shot files of an
    earlier save in ``path`` (``shot_<i>`` / ``ts_<i>`` in any of the formats)
    are removed first, so a load finds this save only; the reference removes
    its fixed ``./shots_time`` directory whole, which here would also take
    other files of a directory the caller names."""
    if save_method not in _EXT:
        raise TypeError("Unknown save_method")
    if save_method == "h5py":
        raise ImportError("h5py is not installed please install it and try again")
    if os.path.isdir(path):
        stale = re.compile(r"^(shot|ts)_[0-9]+(" + "|".join(re.escape(e) for e in _EXT.values()) + r")$")
        for name in os.listdir(path):
            if stale.match(name) and os.path.isfile(os.path.join(path, name)):
                os.remove(os.path.join(path, name))
    os.makedirs(path, exist_ok=True)
    for i, shot in enumerate(shots):
        fshot, fts = _paths(path, i, save_method)
        data, ts = np.asarray(shot.receivers.data), np.asarray(shot.receivers.ts)
        if save_method == "savemat":
            import scipy.io as sio
            sio.savemat(fshot, mdict={fshot: data})
            sio.savemat(fts, mdict={fts: ts})
        elif save_method == "pickle":
            for f, a in ((fshot, data), (fts, ts)):
                with open(f, "wb") as fh:
                    pickle.dump(a, fh, protocol=pickle.HIGHEST_PROTOCOL)
        else:
            np.save(fshot, data, allow_pickle=False)
            np.save(fts, ts, allow_pickle=False)


def load_shot_data(shots, save_method, path="./shots_time"):
    """Read every shot's traces and time axis back onto its receivers."""
    if save_method not in _EXT:
        raise TypeError("Unknown save_method")
    if save_method == "h5py":
        raise ImportError("h5py is not installed please install it and try again")
    if not os.path.exists(path):
        raise ImportError("There is no shot directory, please relaunch generate_seismic_data with save_method "
                          "argument")
    try:
        for i, shot in enumerate(shots):
            fshot, fts = _paths(path, i, save_method)
            if save_method == "savemat":
                import scipy.io as sio
                ts = sio.loadmat(fts)[fts].flatten()
                data = sio.loadmat(fshot)[fshot]
            elif save_method == "pickle":
                with open(fts, "rb") as fh:
                    ts = np.asarray(_ArrayUnpickler(fh).load())
                with open(fshot, "rb") as fh:
                    data = np.asarray(_ArrayUnpickler(fh).load())
            else:
                ts = np.load(fts, allow_pickle=False)
                data = np.load(fshot, allow_pickle=False)
            shot.receivers.clear_data(len(ts))
            shot.receivers.ts = ts
            shot.receivers.data = data
    except (IOError, KeyError, pickle.UnpicklingError) as e:
        raise IOError("Files are corrupted or generated with different save_method argument") from e


def generate_seismic_data_from_file(shots, solver, verbose=False, save_method=None, path="./shots_time",
                                    **kwargs):
    """Load observed data written by ``generate_seismic_data(...,
    save_method=...)`` (data_modeling.py:180-330, time data)."""
    if solver.supports["equation_dynamics"] != "time":
        raise TypeError("A time solver must be specified.")
    if verbose:
        print("Loading data...")
        t0 = time.time()
    if save_method is not None:
        load_shot_data(shots, save_method, path)
    if verbose:
        print(f"Data loading: {time.time() - t0}s")


def generate_seismic_data(shots, solver, model, verbose=False, frequencies=None, save_method=None,
                          batch_size=None, **kwargs):
    if solver.supports["equation_dynamics"] != "time":
        raise TypeError("A time solver must be specified.")
    if verbose:
        print("Generating data...")
        t0 = time.time()
    tm = TemporalModeling(solver)
    solver.model_parameters = model
    ts = solver.ts()
    step = batch_size or shots_per_batch(solver.mesh.shape(include_bc=True, as_grid=True),
                                         np.dtype(solver.dtype).itemsize)
    for i in range(0, len(shots), step):
        group = shots[i:i + step]
        traces = tm.forward_model_batch(group, model)
        for shot, d in zip(group, traces):
            shot.reset_time_series(ts)
            shot.dt = solver.dt
            shot.trange = solver.trange
            shot.receivers.data = np.asarray(d, dtype=np.float64)
    if verbose:
        dt_ = time.time() - t0
        print(f"Data generation: {dt_}s")
        print(f"Data generation: {dt_ / max(1, len(shots))}s/shot")
    if save_method is not None:
        save_shot_data(shots, save_method, kwargs.get("path", "./shots_time"))


def generate_shot_data_time(shot, solver, model, wavefields=None, wavefields_padded=None, verbose=False,
                            **kwargs):
    tm = TemporalModeling(solver)
    solver.model_parameters = model
    shot.reset_time_series(solver.ts())
    shot.dt = solver.dt
    shot.trange = solver.trange
    rp = ["simdata"] + (["wavefield"] if wavefields is not None else [])
    rv = tm.forward_model(shot, model, 1, return_parameters=rp)
    shot.receivers.data = rv["simdata"]
    if wavefields is not None:
        wavefields[:] = rv["wavefield"]
    if wavefields_padded is not None:
        raise NotImplementedError("padded wavefields are not returned by the device loop")
