"""Constant-density acoustic leapfrog solvers backed by the HIP library.

Mirrors pysit/solvers/constant_density_acoustic/time/scalar/
constant_density_acoustic_time_scalar_{2D,3D}.py (classes ``..._2D_cpp`` /
``..._3D_cpp``, :155-190 and :183-227) and the scalar base (:11-83): the
``supports`` contract, the operator components used by dWaveOp
(``sx, sz, sxPsz, sxsz``; 3D ``sxPsyPsz, sxsyPsxszPsysz, sxsysz``), the
wavefield layout (2D aux ``Phix, Phiz``; 3D ``psi, Phix, Phiy, Phiz``) and
``time_step(solver_data, rhs_k, rhs_kp1)``, which advances ``kp1`` in place
from ``km1``, ``k`` and ``rhs_k``.

Here ``kernel_implementation='hip'``: ``time_step`` runs the step on the GPU
through ``pysit_amd_cda_step{2,3}d_{f32,f64}``.  The whole-shot loops of
``TemporalModeling`` do not call it; they run fused on the device (see
``pysit_amd.modeling``).  There is no CPU kernel in this package.
"""
import ctypes

import numpy as np

from .factory import inherit_dict, SolverFactory
from .solver_base import (ConstantDensityAcousticTimeBase, SolverDataTimeBase, WavefieldVectorBase,
                          build_sigma)
from ..core.domain import Bunch

__all__ = ["ConstantDensityAcousticTimeScalarBase", "ConstantDensityAcousticTimeScalar_2D",
           "ConstantDensityAcousticTimeScalar_3D", "ConstantDensityAcousticTimeScalar_2D_hip",
           "ConstantDensityAcousticTimeScalar_3D_hip", "ConstantDensityAcousticWave"]


@inherit_dict("supports", "_local_support_spec")
class ConstantDensityAcousticTimeScalarBase(ConstantDensityAcousticTimeBase):
    _local_support_spec = {"equation_formulation": "scalar", "temporal_integrator": "leap-frog",
                           "temporal_accuracy_order": 2}
    _SolverData = SolverDataTimeBase

    def __init__(self, mesh, **kwargs):
        self.temporal_accuracy_order = 2
        ConstantDensityAcousticTimeBase.__init__(self, mesh, **kwargs)

    def SolverData(self, *args, **kwargs):
        return self._SolverData(self, **kwargs)

    # ---- device-side constants of the current model --------------------
    def kernel_pml(self):
        """Per axis (left, right) PML arrays as the step kernel receives them
        (``mesh.<axis>.{lbc,rbc}.sigma``, empty for non-PML sides)."""
        out = []
        for i in range(self.mesh.dim):
            p = self.mesh.parameters[i]
            l = p.lbc.sigma if p.lbc.type == "pml" else np.zeros(0)
            r = p.rbc.sigma if p.rbc.type == "pml" else np.zeros(0)
            out.append((np.ascontiguousarray(l, dtype=np.float64), np.ascontiguousarray(r, dtype=np.float64)))
        return out

    def model_window(self):
        """Per axis (off, lo, hi): padded node i reads unpadded model node
        clamp(i - off, lo, hi) -- 'edge' padding."""
        return [(self.mesh.parameters[i].lbc.n, 0, self.mesh.parameters[i].n - 1) for i in range(self.mesh.dim)]

    def unpadded_model(self):
        return self._unpadded_model.reshape(tuple(self.mesh.shape(as_grid=True)))

    def _device_constants(self):
        import torch
        key = (self._model_change_count, self.dtype)
        if getattr(self, "_dev_key", None) != key:
            dev = torch.device("cuda", torch.cuda.current_device())
            C = np.ascontiguousarray(self._mp.C.reshape(-1), dtype=self.dtype)
            self._dev_C = torch.from_numpy(C).to(dev)
            self._dev_pml = [torch.from_numpy(np.ascontiguousarray(a, dtype=self.dtype)).to(dev)
                             for pair in self.kernel_pml() for a in pair]
            self._dev_key = key
        return self._dev_C, self._dev_pml

    def time_step(self, solver_data, rhs_k, rhs_kp1):
        import torch
        from .. import _native as N
        if not torch.cuda.is_available():
            raise N.NativeError("kernel_implementation='hip' needs a ROCm GPU")
        C, pml = self._device_constants()
        dev = C.device
        tdt = C.dtype

        def up(a):
            return torch.from_numpy(np.ascontiguousarray(np.asarray(a).reshape(-1), dtype=self.dtype)).to(dev)

        km1, k, kp1 = solver_data.km1, solver_data.k, solver_data.kp1
        ins = [up(km1.u)] + [up(getattr(k, n)) for n in self._step_aux_in] + [up(k.u), C, up(rhs_k)]
        n = self.mesh.dof(include_bc=True)
        outs = [torch.zeros(n, dtype=tdt, device=dev) for _ in self._step_out]
        args = [self.spatial_accuracy_order] + [ctypes.c_void_p(t.data_ptr()) for t in ins]
        for t in pml:
            args += [ctypes.c_void_p(t.data_ptr() if t.numel() else 0), t.numel()]
        args += [self.dt] + list(self.mesh.deltas) + list(self.mesh.shape(include_bc=True, as_grid=True))
        args += [ctypes.c_void_p(t.data_ptr()) for t in outs]
        args += [ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)]
        suffix = "f64" if self.dtype == np.double else "f32"
        N.call(f"pysit_amd_cda_step{self.mesh.dim}d_{suffix}", *args)
        for name, t in zip(self._step_out, outs):
            setattr(kp1, name, t.cpu().numpy().reshape(-1, 1))


@inherit_dict("supports", "_local_support_spec")
class ConstantDensityAcousticTimeScalar_2D(ConstantDensityAcousticTimeScalarBase):
    _local_support_spec = {"spatial_discretization": "finite-difference", "spatial_dimension": 2,
                           "boundary_conditions": ["pml", "pml-sim", "dirichlet"]}
    _step_aux_in = ("Phix", "Phiz")
    _step_out = ("Phix", "Phiz", "u")

    def __init__(self, mesh, spatial_accuracy_order=4, **kwargs):
        self.operator_components = Bunch()
        self.spatial_accuracy_order = spatial_accuracy_order
        ConstantDensityAcousticTimeScalarBase.__init__(self, mesh, spatial_accuracy_order=spatial_accuracy_order,
                                                       **kwargs)
        if self.mesh.x.lbc.type == "pml" and self.mesh.x.lbc.domain_bc.compact:
            raise NotImplementedError("Compact option is not available for time solvers")

    def _rebuild_operators(self):
        oc = self.operator_components
        if not oc.get("_base_components_built", False):
            oc.sx = build_sigma(self.mesh, self.mesh.x)
            oc.sz = build_sigma(self.mesh, self.mesh.z)
            oc.sxPsz = oc.sx + oc.sz
            oc.sxsz = oc.sx * oc.sz
            oc._base_components_built = True

    class WavefieldVector(WavefieldVectorBase):
        aux_names = ["Phix", "Phiz"]


@inherit_dict("supports", "_local_support_spec")
class ConstantDensityAcousticTimeScalar_3D(ConstantDensityAcousticTimeScalarBase):
    _local_support_spec = {"spatial_discretization": "finite-difference", "spatial_dimension": 3,
                           "boundary_conditions": ["pml", "pml-sim", "dirichlet"]}
    _step_aux_in = ("Phix", "Phiy", "Phiz", "psi")
    _step_out = ("Phix", "Phiy", "Phiz", "psi", "u")

    def __init__(self, mesh, spatial_accuracy_order=4, **kwargs):
        self.operator_components = Bunch()
        self.spatial_accuracy_order = spatial_accuracy_order
        ConstantDensityAcousticTimeScalarBase.__init__(self, mesh, spatial_accuracy_order=spatial_accuracy_order,
                                                       **kwargs)

    def _rebuild_operators(self):
        oc = self.operator_components
        if not oc.get("_base_components_built", False):
            oc.sx = build_sigma(self.mesh, self.mesh.x)
            oc.sy = build_sigma(self.mesh, self.mesh.y)
            oc.sz = build_sigma(self.mesh, self.mesh.z)
            oc.sxPsyPsz = oc.sx + oc.sy + oc.sz
            oc.sxsyPsxszPsysz = oc.sx * oc.sy + oc.sx * oc.sz + oc.sy * oc.sz
            oc.sxsysz = oc.sx * oc.sy * oc.sz
            oc._base_components_built = True

    class WavefieldVector(WavefieldVectorBase):
        aux_names = ["psi", "Phix", "Phiy", "Phiz"]


@inherit_dict("supports", "_local_support_spec")
class ConstantDensityAcousticTimeScalar_2D_hip(ConstantDensityAcousticTimeScalar_2D):
    _local_support_spec = {"kernel_implementation": "hip", "spatial_accuracy_order": [2, 4, 6, 8],
                           "precision": ["single", "double"]}


@inherit_dict("supports", "_local_support_spec")
class ConstantDensityAcousticTimeScalar_3D_hip(ConstantDensityAcousticTimeScalar_3D):
    _local_support_spec = {"kernel_implementation": "hip", "spatial_accuracy_order": [2, 4, 6, 8],
                           "precision": ["single", "double"]}


class ConstantDensityAcousticWaveFactory(SolverFactory):
    supports = {"equation_physics": "constant-density-acoustic", "equation_dynamics": "time"}


ConstantDensityAcousticWave = ConstantDensityAcousticWaveFactory()
ConstantDensityAcousticWave.register(ConstantDensityAcousticTimeScalar_2D_hip)
ConstantDensityAcousticWave.register(ConstantDensityAcousticTimeScalar_3D_hip)
