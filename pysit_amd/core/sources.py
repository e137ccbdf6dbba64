"""Seismic sources.

Mirrors pysit/core/sources.py: ``PointSource`` (:56-145) emits
``intensity * w(t)`` through the adjoint point operator; ``SourceSet``
(:183-247) stacks several.  ``f(t)`` returns the dense unpadded right-hand
side like the reference.  ``injection(mesh, dt, nsteps)`` gives the device
path its entry list and per-step amplitudes instead: entries ordered by
source, then by point entry (the order scipy accumulates S*v in), and
``amp[k, j] = intensity_j * w_j(k dt)``.
"""
import numpy as np

from .point_ops import PointRepresentationBase, MeshRepresentationBase, padded_positions

__all__ = ["SourceBase", "PointSource", "SourceSet"]


class SourceBase(object):
    def __init__(self, *args, **kwargs):
        self.shot = None

    def set_shot(self, shot):
        self.shot = shot

    def reset_time_series(self, ts):
        pass

    @property
    def source_count(self):
        return 1


class PointSource(PointRepresentationBase, SourceBase):
    def __init__(self, mesh, pos, w, intensity=1.0, **kwargs):
        PointRepresentationBase.__init__(self, mesh, pos, **kwargs)
        SourceBase.__init__(self, **kwargs)
        self.w = w
        self.intensity = intensity

    def f(self, t=0.0, nu=None, **kwargs):
        amp = self.intensity * (self.w(t, **kwargs) if nu is None else self.w(nu=nu, **kwargs))
        out = self.adjoint_sampling_operator * amp
        if hasattr(out, "toarray"):
            out = out.toarray()
        return np.asarray(out).reshape(self.mesh.shape())

    def _sources(self):
        return [self]

    def injection(self, dt, nsteps):
        return _injection(self.mesh, self._sources(), dt, nsteps)


class SourceSet(SourceBase, MeshRepresentationBase):
    def __init__(self, mesh, sources, **kwargs):
        self.source_list = sources
        SourceBase.__init__(self, **kwargs)
        MeshRepresentationBase.__init__(self, mesh, **kwargs)
        import scipy.sparse as spsp
        if self._sample_interp_method == "sparse":
            self.sampling_operator = spsp.vstack([s.sampling_operator for s in sources])
            self.adjoint_sampling_operator = spsp.hstack([s.adjoint_sampling_operator for s in sources])
        else:
            self.sampling_operator = np.vstack([s.sampling_operator for s in sources])
            self.adjoint_sampling_operator = np.hstack([s.adjoint_sampling_operator for s in sources])

    @property
    def source_count(self):
        return sum(s.source_count for s in self.source_list)

    def set_shot(self, shot):
        self.shot = shot
        for s in self.source_list:
            s.set_shot(shot)

    def w(self, t=None, nu=None, **kwargs):
        if nu is None:
            v = np.array([s.intensity * s.w(t=t, **kwargs) for s in self.source_list])
        else:
            v = np.array([s.intensity * s.w(nu=nu, **kwargs) for s in self.source_list])
        return v.reshape(-1, 1)

    def f(self, t=0.0, nu=None, **kwargs):
        v = self.w(t=t, **kwargs) if nu is None else self.w(nu=nu, **kwargs)
        out = self.adjoint_sampling_operator * v if self._sample_interp_method == "sparse" \
            else np.dot(self.adjoint_sampling_operator, v)
        return np.asarray(out).reshape(self.mesh.shape())

    def _sources(self):
        return list(self.source_list)

    def injection(self, dt, nsteps):
        return _injection(self.mesh, self._sources(), dt, nsteps)


def _injection(mesh, sources, dt, nsteps):
    """(padded positions, channel, weight) and amp (nsteps, n_sources)."""
    pos, chan, w = [], [], []
    amp = np.zeros((nsteps, len(sources)))
    for j, s in enumerate(sources):
        idx, wt = s.entries()
        pos.append(padded_positions(mesh, idx))
        chan.append(np.full(idx.size, j, dtype=np.int64))
        w.append(wt)
        wav = s.w.amplitudes(dt, nsteps) if hasattr(s.w, "amplitudes") else \
            np.array([s.w(k * dt) for k in range(nsteps)], dtype=np.float64)
        amp[:, j] = s.intensity * wav
    return (np.concatenate(pos), np.concatenate(chan), np.concatenate(w)), amp
