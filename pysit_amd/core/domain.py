"""Physical domains and their boundary conditions.

Mirrors pysit/core/domain.py: ``RectangularDomain`` (one config tuple per
axis, z last; :63-141), the boundary conditions ``Dirichlet`` and ``PML``
(:162-254).  ``PML.evaluate`` returns ``amplitude * f(linspace(0, 1, n))``
reversed for the left side (:236-254), the profile the solver's damping
arrays are cut from.
"""
import numpy as np

__all__ = ["DomainBase", "RectangularDomain", "DomainBC", "Dirichlet", "Neumann", "PML", "Bunch"]

AXES = {1: ("z",), 2: ("x", "z"), 3: ("x", "y", "z")}


class Bunch(dict):
    """A dict whose keys are also attributes (survives deepcopy)."""

    def __getattr__(self, key):
        try:
            return self[key]
        except KeyError:
            raise AttributeError(key)

    def __setattr__(self, key, value):
        self[key] = value

    def __delattr__(self, key):
        try:
            del self[key]
        except KeyError:
            raise AttributeError(key)


class DomainBase(object):
    type = None


class RectangularDomain(DomainBase):
    """Axis-aligned box; each config is ``(lbound, rbound, lbc, rbc[, unit])``.

    ``parameters`` is reachable by axis number, by letter and as attributes
    (``d.parameters[1]``, ``d.parameters['z']``, ``d.z``)."""

    type = "rectangular"

    def __init__(self, *configs):
        if not 1 <= len(configs) <= 3:
            raise ValueError("Domain construction must be between one and three dimensions.")
        self.dim = len(configs)
        self.parameters = dict()
        for i, key in enumerate(AXES[self.dim]):
            cfg = configs[i]
            if len(cfg) == 4:
                left, right, lbc, rbc = cfg
                unit = None
            elif len(cfg) == 5:
                left, right, lbc, rbc, unit = cfg
            else:
                raise ValueError("a domain config is (lbound, rbound, lbc, rbc[, unit])")
            p = Bunch(lbound=float(left), rbound=float(right), lbc=lbc, rbc=rbc, unit=unit)
            p.length = float(right) - float(left)
            self.parameters[i] = p
            self.parameters[key] = p
            setattr(self, key, p)

    def collect(self, prop):
        return tuple(self.parameters[i][prop] for i in range(self.dim))

    def get_lengths(self):
        return self.collect("length")


class DomainBC(object):
    type = None


class Dirichlet(DomainBC):
    """Homogeneous Dirichlet boundary."""
    type = "dirichlet"


class Neumann(DomainBC):
    type = "neumann"

    def __init__(self, *args, **kwargs):
        raise NotImplementedError("Neumann boundary support has been temporarily dropped.")


class PML(DomainBC):
    """Perfectly matched layer of physical ``length`` and ``amplitude``.

    ``ftype`` is ``'quadratic'`` (f(x) = x^2) or ``'b-spline'``."""

    type = "pml"

    def __init__(self, length, amplitude, ftype="quadratic", boundary="dirichlet", compact=False):
        self.length = length
        self.amplitude = amplitude
        self.compact = compact
        self.ftype = ftype
        if ftype == "quadratic":
            self.pml_func = self._quadratic
        elif ftype == "b-spline":
            self.pml_func = self._bspline
        else:
            raise NotImplementedError(f"{ftype} is not a currently defined PML function.")
        if boundary not in ("neumann", "dirichlet"):
            raise ValueError(f"'{boundary}' is not 'neumann' or 'dirichlet'.")
        self.boundary_type = boundary

    @staticmethod
    def _quadratic(x):
        return x ** 2

    @staticmethod
    def _bspline(x):
        # cubic B-spline ramp, scaled so that f(1) = 1
        x = np.atleast_1d(np.asarray(x, dtype=np.float64))
        out = np.empty_like(x)
        lo = x < 0.5
        out[lo] = 1.5 * (8.0 / 6.0) * x[lo] ** 3
        hi = ~lo
        out[hi] = 1.5 * (-4.0 * x[hi] ** 3 + 8.0 * x[hi] ** 2 - 4.0 * x[hi] + 2.0 / 3.0)
        return out

    def evaluate(self, n, orientation="right"):
        """The profile on ``n`` points of [0, 1], increasing towards the outer
        edge: reversed for ``orientation='left'``."""
        val = self.amplitude * self.pml_func(np.linspace(0.0, 1.0, n))
        if orientation == "left":
            val = val[::-1]
        return val
