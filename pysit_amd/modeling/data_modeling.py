"""Observed-data generation for time solvers.

Mirrors the time path of pysit/modeling/data_modeling.py
(``generate_seismic_data`` :17-178 and ``generate_shot_data_time``
:335-429): the solver takes the model, every shot's receivers get a fresh
time axis (``reset_time_series(solver.ts())``) and the traces recorded
before each step, and ``shot.dt`` / ``shot.trange`` are set.  All shots run
in one fused device loop.
"""
import time

import numpy as np

from .temporal_modeling import TemporalModeling

__all__ = ["generate_seismic_data", "generate_shot_data_time"]


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
    step = batch_size or len(shots)
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
