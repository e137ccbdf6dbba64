"""Point sources and receivers on a Cartesian mesh.

Mirrors pysit/core/mesh_representation.py:91-212.  A point at ``pos`` is
represented on the unpadded mesh either by a Gaussian (default) or by a
single node (``'delta'``):

* Gaussian: standard deviation ``sigma = max(deltas) * approximation_width``;
  a window of ``ceil(round(deviations * sigma / delta, 3))`` nodes on each side
  of the nearest node ``round((pos - lbound) / delta)``, clipped to the mesh;
  weights ``exp(-|x - pos|^2 / (2 sigma^2)) / ((sqrt(2 pi) sigma)^dim)``.
* delta: weight ``1 / prod(deltas)`` at the nearest node.

The weights integrate to one (``sum(w) * prod(deltas) ~ 1``).  The sampling
operator is ``w * prod(deltas)`` (S), its adjoint is ``w`` (S*) -- the
reference's convention, which makes migrated images L2 gradients.

Besides the scipy operators (kept for API use) every point exposes its
entries -- unpadded flat indices and weights in the reference's accumulation
order -- which the device path turns into injection / sampling tables.
"""
import copy

import numpy as np
import scipy.sparse as spsp

__all__ = ["MeshRepresentationBase", "PointRepresentationBase"]


_SCALARS = (type(None), bool, int, float, complex, str, np.generic)


def _immutable(v):
    """Values a deep copy would return unchanged: scalars, strings, None,
    and tuples of them."""
    return isinstance(v, _SCALARS) or (type(v) is tuple and all(isinstance(x, _SCALARS) for x in v))


class MeshRepresentationBase(object):
    # fixed once built (the mesh and the sampling operators and entries):
    # copies share them, as a deep copy would reproduce them unchanged
    _shared_on_copy = frozenset(("mesh", "domain", "deltas", "sampling_operator", "adjoint_sampling_operator",
                                 "_sampling_operator_base", "_entry_index", "_entry_weight", "position"))

    def __deepcopy__(self, memo):
        new = self.__class__.__new__(self.__class__)
        memo[id(self)] = new
        for k, v in self.__dict__.items():
            if k in self._shared_on_copy or _immutable(v):
                new.__dict__[k] = v
            else:
                new.__dict__[k] = copy.deepcopy(v, memo)
        return new

    def __init__(self, mesh, **kwargs):
        self.mesh = mesh
        self.domain = mesh.domain
        self.deltas = mesh.deltas
        self._prod_deltas = np.prod(self.deltas)
        self._max_deltas = np.max(self.deltas)
        self.sampling_operator = None
        self.adjoint_sampling_operator = None
        self._sample_interp_method = "dense" if self.domain.dim == 1 else "sparse"


class PointRepresentationBase(MeshRepresentationBase):
    def __init__(self, mesh, pos, approximation="gaussian", approximation_width=1,
                 approximation_deviations=3, **kwargs):
        MeshRepresentationBase.__init__(self, mesh, **kwargs)
        if self.domain.dim == 1 and not isinstance(pos, tuple):
            pos = (pos,)
        self.position = pos
        self.approximation = approximation
        self.approximation_width = approximation_width
        self.approximation_deviations = approximation_deviations
        shape = tuple(self.mesh.shape(as_grid=True))
        mins = self.domain.collect("lbound")
        deltas = self.deltas
        # nearest node, Python's round (half to even) as the reference uses
        near = [int(round((p - lo) / d)) for lo, d, p in zip(mins, deltas, pos)]
        if approximation == "gaussian":
            sigma = self._max_deltas * approximation_width
            half = np.ceil(np.around(approximation_deviations * sigma / np.array(deltas), 3)).astype(int)
            coords = self.mesh.mesh_coords(sparse=True)
            ranges = [list(range(max(0, c - w), min(n, c + w + 1))) for c, w, n in zip(near, half, shape)]
            windows = [np.asarray(g).reshape(-1)[r] for g, r in zip(coords, ranges)]
            dim = self.domain.dim
            if dim == 1:
                idx_grids = (np.array(ranges[0]),)
                pos_grids = (np.array(windows[0]),)
            else:
                # 'xy' meshgrid: the entry order of the reference's operator rows
                idx_grids = np.meshgrid(*ranges)
                pos_grids = np.meshgrid(*windows)
            norm = 1.0 / ((np.sqrt(2 * np.pi) ** dim) * (sigma ** dim))
            r2 = None
            for g, p in zip(pos_grids, pos):
                term = (g - p) ** 2
                r2 = term if r2 is None else r2 + term
            weights = (norm * np.exp((-0.5 / sigma ** 2) * r2)).reshape(-1)
            flat = np.ravel_multi_index(np.array([g.reshape(-1) for g in idx_grids]).astype(int), shape)
        elif approximation == "delta":
            flat = np.array([np.ravel_multi_index(tuple(near), shape)])
            weights = np.array([1.0 / self._prod_deltas])
        else:
            raise ValueError("Only 'delta' and 'gaussian' are valid approximations for a point source or receiver.")
        self._entry_index = np.asarray(flat, dtype=np.int64)
        self._entry_weight = np.asarray(weights, dtype=np.float64)
        dof = self.mesh.dof()
        base = spsp.csr_matrix((self._entry_weight, self._entry_index,
                                np.array([0, self._entry_index.size])), shape=(1, dof))
        if self._sample_interp_method == "dense":
            base = base.toarray()
        self._sampling_operator_base = base
        self.sampling_operator = base * self._prod_deltas
        self.adjoint_sampling_operator = base.T

    def entries(self):
        """(unpadded flat indices, weights) of the representation, in order."""
        return self._entry_index, self._entry_weight


def padded_positions(mesh, flat_unpadded):
    """Padded grid coordinates (n, dim) of unpadded flat indices."""
    sh = tuple(mesh.shape(as_grid=True))
    coords = np.stack(np.unravel_index(np.asarray(flat_unpadded, dtype=np.int64), sh), axis=1)
    pads = np.array([mesh.parameters[i].lbc.n for i in range(mesh.dim)])
    return coords + pads[None, :]
