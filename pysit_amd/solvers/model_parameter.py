"""Model parameters and perturbations.

Mirrors pysit/solvers/model_parameter.py: a model stores its parameters in
nonlinear form (the wave speed C) and does all arithmetic in linear form
(m = C^-2, ``WaveSpeed.linearize``, :80-100); the difference of two models
and every gradient are ``Perturbation``s, always linear.  Perturbations carry
the mesh inner product (prod(deltas) weighted, :799-808) and ``add_padding``,
the fold of a padded image onto the physical grid (:537-593; the transpose of
'edge' padding).
"""
import copy
import itertools

import numpy as np

__all__ = ["ModelParameterBase", "ModelPerturbationBase", "WaveSpeed", "ConstantDensityAcousticParameters",
           "fold_padding"]


def _finite(x):
    return np.all(np.isfinite(x))


def _positive(x):
    return np.all(x > 0)


class ModelParameterDescription(object):
    name = None
    constraints = tuple()
    postprocessing_steps = []

    @classmethod
    def validate(cls, data):
        for f, args, label in cls.constraints:
            if not f(data, *args):
                return False, label
        return True, None

    @classmethod
    def postprocess(cls, data):
        for f, args in cls.postprocessing_steps:
            f(data, *args)


def _cap_above(x, v):
    x[x >= v] = v


def _cap_below(x, v):
    x[x <= v] = v


class WaveSpeed(ModelParameterDescription):
    name = "C"
    constraints = ((_finite, tuple(), "finite"), (_positive, tuple(), "positivity"))
    postprocessing_steps = []

    @classmethod
    def linearize(cls, data):
        return data ** -2

    @classmethod
    def unlinearize(cls, data):
        return np.sqrt(1.0 / data)

    @classmethod
    def add_upper_bound(cls, val):
        cls.postprocessing_steps.append((_cap_above, (val,)))

    @classmethod
    def add_lower_bound(cls, val):
        cls.postprocessing_steps.append((_cap_below, (val,)))


def fold_padding(mesh, data_padded):
    """Sum each PML slab onto the first / last physical plane (x, then y,
    then z) and crop: the unpadded array (host, FP64)."""
    sh = tuple(mesh.shape(include_bc=True, as_grid=True))
    a = np.array(data_padded, dtype=np.result_type(data_padded, np.float64)).reshape(sh)
    for ax in range(mesh.dim):
        l = mesh.parameters[ax].lbc.n
        r = mesh.parameters[ax].rbc.n
        n_in = sh[ax] - l - r
        a = np.moveaxis(a, ax, 0)
        if l:
            a[l] = a[l] + a[:l].sum(axis=0)
        if r:
            a[l + n_in - 1] = a[l + n_in - 1] + a[l + n_in:].sum(axis=0)
        a = np.moveaxis(a, 0, ax)
    sl = tuple(slice(mesh.parameters[i].lbc.n, sh[i] - mesh.parameters[i].rbc.n) for i in range(mesh.dim))
    return a[sl].reshape(-1, 1)


class _Base(object):
    parameter_list = []

    def _dof(self):
        return self.mesh.dof(include_bc=self.padded)

    def _slices(self):
        n = self._dof()
        return [(p, slice(i * n, (i + 1) * n)) for i, p in enumerate(self.parameter_list)]

    def M(self, i):
        if type(i) is not int:
            raise TypeError("__call__ method used for indexing the list of models")
        if not 0 <= i < len(self.parameter_list):
            raise IndexError("Parameter index out of bounds.")
        n = self._dof()
        return self.data[i * n:(i + 1) * n, 0].reshape(-1, 1)

    def asarray(self):
        return self.data

    def norm(self, ord=None):
        return np.linalg.norm(self.data, ord=ord)

    @property
    def T(self):
        return self.data.T

    def __repr__(self):
        return repr(self.data)


def _add_param_properties(cls):
    for i, p in enumerate(cls.parameter_list):
        def getter(self, i=i):
            n = self._dof()
            return self.data[i * n:(i + 1) * n]

        def setter(self, value, i=i):
            n = self._dof()
            self.data[i * n:(i + 1) * n] = value
        setattr(cls, p.name, property(getter, setter))
    return cls


class ModelParameterBase(_Base):
    """Container of nonlinear model parameters (``data`` is (dof, 1))."""

    class Perturbation(object):
        pass

    def __init__(self, mesh, inputs=None, linear_inputs=None, padded=False):
        self.padded = padded
        self.mesh = mesh
        n = self._dof()
        self.parameter_count = len(self.parameter_list)
        self.data = np.ones((self.parameter_count * n, 1)) * np.inf
        if inputs is not None:
            if isinstance(inputs, (tuple, list)):
                for (p, sl), v in zip(self._slices(), inputs):
                    self.data[sl] = 0
                    self.data[sl] += v
            elif isinstance(inputs, dict):
                for p, sl in self._slices():
                    if p.name in inputs:
                        self.data[sl] = 0
                        self.data[sl] += inputs[p.name]
            elif isinstance(inputs, np.ndarray):
                if inputs.size != self.data.size:
                    raise ValueError("Improper dimensions for input array.")
                self.data = inputs
                self.data.shape = -1, 1
            else:
                raise ValueError("Invalid format for collection of input parameters.")
        if linear_inputs is not None:
            arr = linear_inputs if isinstance(linear_inputs, (tuple, list)) else None
            for k, (p, sl) in enumerate(self._slices()):
                v = arr[k] if arr is not None else np.asarray(linear_inputs).reshape(-1, 1)[sl]
                self.data[sl] = 0
                self.data[sl] += p.unlinearize(v)
        for p, sl in self._slices():
            p.postprocess(self.data[sl])

    def perturbation(self, data=None, *args, **kwargs):
        if data is None:
            return self.Perturbation(self.mesh, padded=self.padded, *args, **kwargs)
        return self.Perturbation(self.mesh, padded=self.padded, inputs=data, *args, **kwargs)

    def linearize(self, asperturbation=False):
        out = np.zeros_like(self.data, dtype=np.float64)
        for p, sl in self._slices():
            out[sl] += p.linearize(self.data[sl])
        return self.perturbation(out) if asperturbation else out

    def validate(self, raise_exception=False):
        ok = True
        for p in self.parameter_list:
            good, label = p.validate(getattr(self, p.name))
            if not good:
                ok = False
                if raise_exception:
                    raise Exception(f"Validation of model failed: {label}")
        return ok

    def with_padding(self, **kwargs):
        if self.padded:
            return self
        out = type(self)(self.mesh, padded=True)
        nn = self.mesh.dof(include_bc=True)
        no = self.mesh.dof(include_bc=False)
        for i in range(len(self.parameter_list)):
            out.data[i * nn:(i + 1) * nn] = 0
            out.data[i * nn:(i + 1) * nn] += self.mesh.pad_array(self.data[i * no:(i + 1) * no], **kwargs)
        return out

    def without_padding(self):
        if not self.padded:
            return self
        out = type(self)(self.mesh, padded=False)
        nn = self.mesh.dof(include_bc=True)
        no = self.mesh.dof(include_bc=False)
        for i in range(len(self.parameter_list)):
            out.data[i * no:(i + 1) * no] = 0
            out.data[i * no:(i + 1) * no] += self.mesh.unpad_array(self.data[i * nn:(i + 1) * nn])
        return out

    def _new(self):
        return type(self)(self.mesh, padded=self.padded)

    def _finish(self, result):
        for p, sl in result._slices():
            p.postprocess(result.data[sl])
        return result

    def _combine(self, rhs, op):
        """unlinearize(op(linearize(self), rhs_linear)) per parameter."""
        out = self._new()
        for k, (p, sl) in enumerate(self._slices()):
            lin = p.linearize(self.data[sl])
            if isinstance(rhs, (list, tuple)) or (isinstance(rhs, np.ndarray) and rhs.ndim == 1
                                                  and len(rhs) == self.parameter_count):
                other = rhs[k]
            elif type(rhs) is type(self) and rhs.data.shape == self.data.shape:
                other = p.linearize(rhs.data[sl])
            elif isinstance(rhs, np.ndarray) and rhs.shape == self.data.shape:
                other = rhs[sl]
            elif isinstance(rhs, ModelPerturbationBase) and rhs.data.shape == self.data.shape:
                other = rhs.data[sl]
            else:
                other = rhs
            out.data[sl] = 0
            out.data[sl] += p.unlinearize(op(lin, other))
        return self._finish(out)

    def __add__(self, rhs):
        return self._combine(rhs, lambda a, b: a + b)

    __radd__ = __add__

    def __mul__(self, rhs):
        if isinstance(rhs, ModelPerturbationBase):
            return self.linearize() * rhs.data
        if isinstance(rhs, np.ndarray) and rhs.shape == self.data.shape:
            return self.linearize() * rhs
        return self._combine(rhs, lambda a, b: b * a)

    __rmul__ = __mul__

    def __sub__(self, rhs):
        if (type(rhs) is type(self) or isinstance(rhs, ModelPerturbationBase)) and rhs.data.shape == self.data.shape:
            arr = np.zeros_like(self.data, dtype=np.float64)
            for p, sl in self._slices():
                other = rhs.data[sl] if isinstance(rhs, ModelPerturbationBase) else p.linearize(rhs.data[sl])
                arr[sl] += p.linearize(self.data[sl]) - other
            return self.perturbation(data=arr)
        if isinstance(rhs, np.ndarray) and rhs.shape == self.data.shape:
            return self._combine(rhs, lambda a, b: a - b)
        if isinstance(rhs, (list, tuple)):
            return self._combine(rhs, lambda a, b: a - b)
        arr = np.zeros_like(self.data, dtype=np.float64)
        for p, sl in self._slices():
            arr[sl] = p.linearize(self.data[sl]) - rhs
        return self.perturbation(data=arr)

    def __deepcopy__(self, memo):
        return type(self)(self.mesh, padded=self.padded, inputs=copy.deepcopy(self.data, memo))


class ModelPerturbationBase(_Base):
    """Linear model perturbation (a gradient, a step)."""

    def __init__(self, mesh, padded=False, inputs=None, **kwargs):
        self.mesh = mesh
        self.padded = padded
        n = self._dof()
        self.parameter_count = len(self.parameter_list)
        self.dtype = kwargs.get("dtype", np.double)
        self.data = np.zeros((self.parameter_count * n, 1), dtype=self.dtype)
        if inputs is not None:
            if isinstance(inputs, (tuple, list)):
                for (p, sl), v in zip(self._slices(), inputs):
                    self.data[sl] = 0
                    self.data[sl] += v
            elif isinstance(inputs, np.ndarray):
                if inputs.size != self.data.size:
                    raise ValueError("Improper dimensions for input array.")
                self.data[:] = inputs.reshape(self.data.shape)
            else:
                raise ValueError("Invalid format for collection of input parameters.")

    def _new(self):
        return type(self)(self.mesh, padded=self.padded, dtype=self.dtype)

    def with_padding(self, **kwargs):
        if self.padded:
            return self
        out = type(self)(self.mesh, padded=True, dtype=self.dtype)
        out.data[:] = self.mesh.pad_array(self.data, **kwargs).reshape(out.data.shape)
        return out

    def without_padding(self):
        if not self.padded:
            return self
        out = type(self)(self.mesh, padded=False, dtype=self.dtype)
        out.data[:] = self.mesh.unpad_array(self.data).reshape(out.data.shape)
        return out

    def add_padding(self):
        """Fold the padded perturbation onto the physical grid."""
        out = type(self)(self.mesh, padded=False, dtype=self.dtype)
        out.data = fold_padding(self.mesh, self.data)
        return out

    def _other(self, rhs, sl, k):
        if isinstance(rhs, (list, tuple)):
            return rhs[k]
        if isinstance(rhs, np.ndarray) and rhs.shape == self.data.shape:
            return rhs[sl]
        if type(rhs) is type(self) and rhs.data.shape == self.data.shape:
            return rhs.data[sl]
        return rhs

    def _binary(self, rhs, op):
        out = self._new()
        for k, (p, sl) in enumerate(self._slices()):
            out.data[sl] = op(self.data[sl], self._other(rhs, sl, k))
        return out

    def __add__(self, rhs):
        if isinstance(rhs, ModelParameterBase):
            return rhs.__add__(self)
        return self._binary(rhs, lambda a, b: a + b)

    __radd__ = __add__

    def __sub__(self, rhs):
        return self._binary(rhs, lambda a, b: a - b)

    def __mul__(self, rhs):
        if (isinstance(rhs, np.ndarray) and rhs.shape == self.data.shape):
            return self.data * rhs
        if type(rhs) is type(self):
            return self.data * rhs.data
        return self._binary(rhs, lambda a, b: b * a)

    __rmul__ = __mul__

    def __neg__(self):
        return self * -1

    def __iadd__(self, rhs):
        for k, (p, sl) in enumerate(self._slices()):
            self.data[sl] += self._other(rhs, sl, k)
        return self

    def __isub__(self, rhs):
        for k, (p, sl) in enumerate(self._slices()):
            self.data[sl] -= self._other(rhs, sl, k)
        return self

    def __imul__(self, rhs):
        if type(rhs) is type(self):
            self.data *= rhs.data
        elif isinstance(rhs, np.ndarray) and rhs.shape == self.data.shape:
            self.data *= rhs
        else:
            for k, (p, sl) in enumerate(self._slices()):
                self.data[sl] *= self._other(rhs, sl, k)
        return self

    def toreal(self):
        if np.iscomplexobj(self.data):
            self.data = np.real(self.data)
            self.dtype = self.data.dtype

    def inner_product(self, other):
        if type(other) is not type(self):
            raise ValueError("Perturbation inner product is only defined for perturbations.")
        ip = 0.0
        for p, sl in self._slices():
            ip += self.mesh.inner_product(self.data[sl], other.data[sl])
        return ip

    def __deepcopy__(self, memo):
        return type(self)(self.mesh, inputs=copy.deepcopy(self.data, memo), padded=self.padded, dtype=self.dtype)


class ConstantDensityAcousticParameters(ModelParameterBase):
    parameter_list = [WaveSpeed]

    class Perturbation(ModelPerturbationBase):
        parameter_list = [WaveSpeed]


_add_param_properties(ConstantDensityAcousticParameters)
_add_param_properties(ConstantDensityAcousticParameters.Perturbation)
