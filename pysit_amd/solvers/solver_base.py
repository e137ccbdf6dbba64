"""Solver base classes, wavefield containers and per-step solver data.

Mirrors pysit/solvers/solver_base.py:22-135 (``supports`` matching, the
``model_parameters`` setter that edge-pads the model and resets dt),
solver_data.py (the three-level time state ``kp1``/``k``/``km1`` rotated by
``advance``), wavefield_vector.py:7-146 (u plus auxiliary fields in one
(dof*(1+n_aux), 1) array), constant_density_acoustic_base.py:24-40 (the
imaging-condition factor dWaveOp) and
constant_density_acoustic/time/constant_density_acoustic_time_base.py:30-76
(dt = cfl * min(deltas) / max|C|, nsteps = ceil((tf - t0) / dt)).
"""
import math

import numpy as np

from .factory import supports, inherit_dict, DoesNotMatch, IncompleteMatch, CompleteMatch
from .model_parameter import ConstantDensityAcousticParameters

__all__ = ["SolverBase", "WavefieldVectorBase", "SolverDataTimeBase", "ConstantDensityAcousticBase",
           "ConstantDensityAcousticTimeBase", "build_sigma"]


def build_sigma(mesh, axis_params):
    """Mesh PML profile of one axis over the padded grid (flat): the sum of
    both sides' ``eval_on_mesh`` (pysit/util/matrix_helpers.py:6-13)."""
    s = np.zeros(mesh.shape(include_bc=True))
    if axis_params.lbc.type == "pml":
        s += axis_params.lbc.eval_on_mesh().reshape(s.shape)
    if axis_params.rbc.type == "pml":
        s += axis_params.rbc.eval_on_mesh().reshape(s.shape)
    return s.reshape(-1)


class WavefieldVectorBase(object):
    aux_names = []

    def __init__(self, mesh, dtype=np.double, data=None):
        self.mesh = mesh
        self.n_aux = len(self.aux_names)
        self.n_fields = 1 + self.n_aux
        self.dof = mesh.dof(include_bc=True)
        self.data_shape = (self.dof * self.n_fields, 1)
        if data is None:
            self.dtype = dtype
            self._data = np.zeros(self.data_shape, dtype=dtype)
        else:
            if data.size != self.n_fields * self.dof:
                raise ValueError("Size mismatch between mesh and proposed data.")
            self.dtype = data.dtype
            self._data = data
            self._data.shape = self.data_shape

    @property
    def data(self):
        return self._data

    @data.setter
    def data(self, arg):
        self._data[:] = arg

    def _field(self, i):
        return self._data[i * self.dof:(i + 1) * self.dof]

    @property
    def u(self):
        return self._field(0)

    @u.setter
    def u(self, v):
        self._data[:self.dof] = v

    primary_wavefield = u

    def __getattr__(self, name):
        names = type(self).aux_names
        if name in names:
            return self._field(names.index(name) + 1)
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in type(self).aux_names:
            i = type(self).aux_names.index(name) + 1
            self._data[i * self.dof:(i + 1) * self.dof] = value
        else:
            object.__setattr__(self, name, value)

    def _binary(self, rhs, op):
        out = type(self)(self.mesh, dtype=self.dtype)
        out.data = op(self.data, rhs.data if type(rhs) is type(self) else rhs)
        return out

    def __add__(self, rhs):
        return self._binary(rhs, lambda a, b: a + b)

    __radd__ = __add__

    def __sub__(self, rhs):
        return self._binary(rhs, lambda a, b: a - b)

    def __mul__(self, rhs):
        return self._binary(rhs, lambda a, b: a * b)

    __rmul__ = __mul__

    def __iadd__(self, rhs):
        self.data += rhs.data if type(rhs) is type(self) else rhs
        return self

    def __isub__(self, rhs):
        self.data -= rhs.data if type(rhs) is type(self) else rhs
        return self

    def __imul__(self, rhs):
        self.data *= rhs.data if type(rhs) is type(self) else rhs
        return self

    def __deepcopy__(self, memo):
        return type(self)(self.mesh, dtype=self.dtype, data=self._data.copy())


class SolverDataTimeBase(object):
    """us[0] = kp1, us[1] = k, us[2] = km1; ``advance`` zeroes km1 and rotates."""

    def __init__(self, solver, **kwargs):
        self.solver = solver
        self.us = [solver.WavefieldVector(solver.mesh, dtype=solver.dtype) for _ in range(3)]

    def advance(self):
        self.us[-1] *= 0
        self.us.insert(0, self.us.pop(-1))

    kp1 = property(lambda self: self.us[0], lambda self, v: self.us.__setitem__(0, v))
    k = property(lambda self: self.us[1], lambda self, v: self.us.__setitem__(1, v))
    km1 = property(lambda self: self.us[2], lambda self, v: self.us.__setitem__(2, v))


@inherit_dict("supports", "_local_support_spec")
class SolverBase(object):
    _local_support_spec = {"equation_physics": None, "equation_dynamics": None}
    ModelParameters = None
    WavefieldVector = None

    def __init__(self, mesh, precision="double", **kwargs):
        self.mesh = mesh
        self.domain = mesh.domain
        if precision in self.supports["precision"]:
            self.precision = precision
            if self.supports["equation_dynamics"] == "time":
                self.dtype = np.double if precision == "double" else np.single
            else:
                self.dtype = np.complex128 if precision == "double" else np.complex64
        else:
            raise ValueError(f"precision {precision!r} is not supported")
        self._mp = None
        self._model_change_count = 0

    @classmethod
    def _factory_validation_function(cls, mesh, *args, **kwargs):
        complete = True
        for parameter, values in cls.supports.items():
            if parameter in kwargs:
                if not supports(kwargs[parameter], values):
                    return DoesNotMatch
            elif parameter == "boundary_conditions":
                for i in range(mesh.dim):
                    if not (supports(mesh.parameters[i].lbc.type, values)
                            and supports(mesh.parameters[i].rbc.type, values)):
                        return DoesNotMatch
            elif parameter == "spatial_dimension":
                if not supports(mesh.dim, values):
                    return DoesNotMatch
            else:
                complete = False
        return CompleteMatch if complete else IncompleteMatch

    @property
    def model_parameters(self):
        return self._mp

    @model_parameters.setter
    def model_parameters(self, mp):
        if type(mp) is not self.ModelParameters:
            raise TypeError(f"{type(mp)} is not of type {self.ModelParameters}")
        if self._mp is None or np.linalg.norm(self._mp.without_padding().data - mp.data) != 0.0:
            self._mp = mp.with_padding(padding_mode="edge")
            self._unpadded_model = np.array(mp.data, dtype=np.float64)
            self._process_mp_reset()
            self._model_change_count += 1

    def _process_mp_reset(self, *args, **kwargs):
        raise NotImplementedError("_process_mp_reset() must be implemented by a subclass.")

    def compute_dWaveOp(self, regime, *args):
        return getattr(self, f"_compute_dWaveOp_{regime}")(*args)

    def SolverData(self, **kwargs):
        return self._SolverData(self, **kwargs)


@inherit_dict("supports", "_local_support_spec")
class ConstantDensityAcousticBase(SolverBase):
    _local_support_spec = {"equation_physics": "constant-density-acoustic"}
    ModelParameters = ConstantDensityAcousticParameters

    def _compute_dWaveOp_time(self, solver_data):
        """(u+ - 2u + u-)/dt^2 + mesh-PML terms, on the padded grid (host)."""
        up = solver_data.kp1.primary_wavefield
        u = solver_data.k.primary_wavefield
        um = solver_data.km1.primary_wavefield
        oc = self.operator_components
        dt = self.dt
        if self.mesh.dim == 2:
            pml = oc.sxPsz.reshape(-1, 1) * (up - um) / (2.0 * dt) + oc.sxsz.reshape(-1, 1) * u
        else:
            pml = (oc.sxPsyPsz.reshape(-1, 1) * (up - um) / (2.0 * dt)
                   + oc.sxsyPsxszPsysz.reshape(-1, 1) * u + oc.sxsysz.reshape(-1, 1) * solver_data.k.psi)
        return (up - 2 * u + um) / (dt ** 2) + pml


@inherit_dict("supports", "_local_support_spec")
class ConstantDensityAcousticTimeBase(ConstantDensityAcousticBase):
    _local_support_spec = {"equation_dynamics": "time", "temporal_integrator": None,
                           "temporal_accuracy_order": None, "spatial_discretization": None,
                           "spatial_accuracy_order": None, "kernel_implementation": None,
                           "spatial_dimension": None, "boundary_conditions": None, "precision": None}

    def __init__(self, mesh, trange=(0.0, 1.0), cfl_safety=1 / 6, **kwargs):
        self.trange = trange
        self.cfl_safety = cfl_safety
        self.t0, self.tf = trange
        self.dt = 0.0
        self.nsteps = 0
        ConstantDensityAcousticBase.__init__(self, mesh, **kwargs)

    def ts(self):
        return np.arange(self.nsteps) * self.dt

    def _process_mp_reset(self, *args, **kwargs):
        t0, tf = self.trange
        C = self._mp.C
        max_c = max(abs(C.min()), C.max())
        dt = self.cfl_safety * np.min(self.mesh.deltas) / max_c
        self.dt = dt
        self.nsteps = int(math.ceil((tf - t0) / dt))
        self._rebuild_operators()

    def _rebuild_operators(self, *args, **kwargs):
        pass
