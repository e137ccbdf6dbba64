"""Cartesian meshes with boundary padding.

Mirrors pysit/core/mesh.py: ``CartesianMesh`` (:104-538) with ``shape``,
``dof``, ``deltas``, ``mesh_coords``, ``pad_array`` / ``unpad_array`` and the
mesh inner product; the boundary descriptions ``StructuredDirichlet`` and
``StructuredPML`` (:657-830).

The PML of ``n = ceil(length / delta)`` nodes takes its damping profile from
``PML.evaluate(n + 1, side)`` with the node on the physical boundary dropped:
``sigma = s[:-1]`` on the left, ``s[1:]`` on the right (:749-782).  Those two
arrays are exactly what the step kernel receives.  ``eval_on_mesh`` spreads
one side's profile over the padded grid (:784-830).
"""
import numpy as np

from .domain import AXES, Bunch

__all__ = ["CartesianMesh", "StructuredDirichlet", "StructuredPML", "MeshBC"]


class StructuredBCBase(object):
    type = None

    def __init__(self, mesh, domain_bc, dim, side, delta):
        self.mesh = mesh
        self.domain_bc = domain_bc
        self._n = 0
        self.solver_padding = 1
        self.dim = dim
        self.side = side

    @property
    def n(self):
        """Padding nodes on this side."""
        return self._n

    def __getitem__(self, key):  # model_parameter code reads mesh.z['lbc']._n
        return getattr(self, key)


class StructuredDirichlet(StructuredBCBase):
    type = "dirichlet"
    boundary_type = "dirichlet"


class StructuredPML(StructuredBCBase):
    type = "pml"

    def __init__(self, mesh, domain_bc, dim, side, delta):
        StructuredBCBase.__init__(self, mesh, domain_bc, dim, side, delta)
        self._n = int(np.ceil(domain_bc.length / delta))
        s = domain_bc.evaluate(self._n + 1, side)
        self.sigma = np.ascontiguousarray(s[1:] if side == "right" else s[:-1])
        self.boundary_type = domain_bc.boundary_type

    def eval_on_mesh(self):
        """This side's profile laid on the padded grid (flat, padded dof)."""
        shape = tuple(self.mesh.shape(include_bc=True, as_grid=True))
        out = np.zeros(shape)
        sl = [slice(None)] * self.mesh.dim
        sl[self.dim] = slice(0, self._n) if self.side == "left" else slice(shape[self.dim] - self._n, None)
        bshape = [1] * self.mesh.dim
        bshape[self.dim] = -1
        out[tuple(sl)] = self.sigma.reshape(bshape)
        return out.reshape(-1)


def MeshBC(mesh, domain_bc, dim, side, delta):
    """Mesh-level boundary for a domain boundary condition."""
    if domain_bc.type == "dirichlet":
        return StructuredDirichlet(mesh, domain_bc, dim, side, delta)
    if domain_bc.type == "pml":
        return StructuredPML(mesh, domain_bc, dim, side, delta)
    raise NotImplementedError(f"boundary type {domain_bc.type!r} is not supported")


class CartesianMesh(object):
    """Uniform mesh of a ``RectangularDomain``; ``configs`` are node counts
    per axis (including both boundary nodes), z last."""

    type = "structured-cartesian"

    def __init__(self, domain, *configs):
        self.domain = domain
        self.dim = domain.dim
        if len(configs) != self.dim:
            raise ValueError("one node count per domain axis")
        self.parameters = dict()
        for i, key in enumerate(AXES[self.dim]):
            n = int(configs[i])
            delta = domain.parameters[i].length / (n - 1)
            p = Bunch(n=n, delta=delta)
            p.lbc = MeshBC(self, domain.parameters[i].lbc, i, "left", delta)
            p.rbc = MeshBC(self, domain.parameters[i].rbc, i, "right", delta)
            self.parameters[i] = p
            self.parameters[key] = p
            setattr(self, key, p)
        self._shapes = dict()
        self._dofs = dict()

    @property
    def deltas(self):
        return tuple(self.parameters[i].delta for i in range(self.dim))

    def _compute_shape(self, include_bc):
        sh = []
        for i in range(self.dim):
            p = self.parameters[i]
            n = p.n + (p.lbc.n + p.rbc.n if include_bc else 0)
            sh.append(n)
        self._shapes[(include_bc, True)] = sh
        self._shapes[(include_bc, False)] = (int(np.prod(sh)), 1)
        self._dofs[include_bc] = int(np.prod(sh))

    def shape(self, include_bc=False, as_grid=False):
        if (include_bc, as_grid) not in self._shapes:
            self._compute_shape(include_bc)
        return self._shapes[(include_bc, as_grid)]

    def dof(self, include_bc=False):
        if include_bc not in self._dofs:
            self._compute_shape(include_bc)
        return self._dofs[include_bc]

    def pad_widths(self):
        return tuple((self.parameters[i].lbc.n, self.parameters[i].rbc.n) for i in range(self.dim))

    def mesh_coords(self, sparse=False, include_bc=False):
        rows = []
        for i in range(self.dim):
            p = self.parameters[i]
            d = self.domain.parameters[i]
            if include_bc:
                rows.append(np.linspace(d.lbound - p.lbc.n * p.delta, d.rbound + p.rbc.n * p.delta,
                                        p.n + p.lbc.n + p.rbc.n))
            else:
                rows.append(np.linspace(d.lbound, d.rbound, p.n))
        if self.dim == 1:
            return (rows[0],)
        grids = np.meshgrid(*rows, sparse=sparse, indexing="ij")
        if not sparse:
            grids = [g.reshape(self.shape(include_bc)) for g in grids]
        return tuple(grids)

    def nodes(self, include_bc=False):
        return np.hstack(self.mesh_coords(include_bc=include_bc))

    def _interior_slices(self):
        sh = self.shape(include_bc=True, as_grid=True)
        return tuple(slice(self.parameters[i].lbc.n, sh[i] - self.parameters[i].rbc.n) for i in range(self.dim))

    def unpad_array(self, in_array, copy=False):
        """The unpadded part of a padded array; keeps grid or (dof, 1) shape."""
        if in_array.shape in (tuple(self.shape(False, True)), self.shape(False, False)):
            out = in_array
        else:
            out = in_array.reshape(self.shape(True, True))[self._interior_slices()]
            if in_array.ndim == 2 and in_array.shape[1] == 1:
                out = out.reshape(-1, 1)
        return out.copy() if copy else out

    def pad_array(self, in_array, out_array=None, padding_mode=None):
        """Pad to the boundary-inclusive grid, with zeros (default) or any
        ``numpy.pad`` mode ('edge' repeats the boundary value)."""
        sh_in_vec = self.shape(False, False)
        sh_in = tuple(self.shape(False, True))
        sh_out = tuple(self.shape(True, True))
        if out_array is not None:
            out_array.shape = sh_out
        if padding_mode is not None:
            padded = np.pad(in_array.reshape(sh_in), self.pad_widths(), mode=padding_mode)
            if out_array is not None:
                out_array[:] = padded
            else:
                out_array = padded
        else:
            if out_array is None:
                out_array = np.zeros(sh_out, dtype=in_array.dtype)
            out_array[self._interior_slices()] = in_array.reshape(sh_in)
        out_array.shape = self.shape(True, False) if in_array.shape == sh_in_vec else sh_out
        return out_array

    def inner_product(self, arg1, arg2):
        """Discrete L2 inner product: sum of products times prod(deltas)."""
        return np.dot(arg1.T, arg2).squeeze() * np.prod(self.deltas)
