"""Seismic receivers.

Mirrors pysit/core/receivers.py: ``PointReceiver`` (:134-335) and
``ReceiverSet`` (:362-624) -- data storage, ``sample_data_from_array``
(S u, weights times prod(deltas)), ``extend_data_to_array`` (S* d, weights
alone), and ``interpolate_data``: linear interpolation of the stored data onto
another time axis with zero outside it (:56-105, scipy ``interp1d``).
``sampling_entries`` / ``adjoint_entries`` give the device path the same
operators as entry lists grouped by receiver.
"""
import numpy as np
from scipy.interpolate import interp1d

from .point_ops import PointRepresentationBase, MeshRepresentationBase, padded_positions

__all__ = ["PointReceiver", "ReceiverSet"]


class ReceiverBase(object):
    def __init__(self, time_window=None, **kwargs):
        self.ts = None
        self._data = None
        # bumped whenever the data change through this class (assignment or
        # sampling into them), so device copies know when to upload again
        self.data_version = 0
        self.shot = None
        self._time_window = ("None",) if time_window is None else time_window

    def set_shot(self, shot):
        self.shot = shot

    @property
    def data(self):
        return self._data

    @data.setter
    def data(self, value):
        self._data = value
        self.data_version = getattr(self, "data_version", 0) + 1

    def reset_time_series(self, ts):
        self.ts = ts
        self.clear_data(len(ts))

    def interpolate_data(self, ts):
        if self.ts is None or self.data is None:
            raise TypeError("Interpolator has not been defined for the current receiver.")
        f = interp1d(self.ts, self.data, axis=0, kind="linear", copy=False, bounds_error=False,
                     fill_value=0.0)
        d = f(ts)
        if self._time_window[0] not in ("None", "Set"):
            d = d * self.time_window(ts)
        return d

    def time_window(self, ts):
        kind = self._time_window[0]
        if kind == "Box":
            lb, rb = self._time_window[1], self._time_window[2]
            if lb is None:
                return 1.0 * (ts < rb)
            if rb is None:
                return 1.0 * (ts > lb)
            return 1.0 * ((ts > lb) & (ts < rb))
        if kind == "Gaussian":
            t0, s = self._time_window[1], self._time_window[2]
            return np.exp(-(ts - t0) ** 2 / s ** 2)
        return 1.0


class PointReceiver(PointRepresentationBase, ReceiverBase):
    def __init__(self, mesh, pos, **kwargs):
        PointRepresentationBase.__init__(self, mesh, pos, **kwargs)
        ReceiverBase.__init__(self, **kwargs)

    @property
    def receiver_count(self):
        return 1

    def clear_data(self, n):
        self.data = np.zeros((n, 1))

    def sample_data_from_array(self, arr, k=None, data=None):
        v = self.sampling_operator * arr if self._sample_interp_method == "sparse" \
            else np.dot(self.sampling_operator, arr)
        v = np.asarray(v).reshape(-1)[0]
        if k is None:
            return v
        if data is None:
            self.data_version = getattr(self, "data_version", 0) + 1
        (self.data if data is None else data)[k] = v

    def extend_data_to_array(self, k=None, resid=False, data=None):
        if k is not None:
            d = self.data[k] if data is None else ((self.data[k] - data[k]) if resid else data[k])
        else:
            d = data
        out = self.adjoint_sampling_operator * d
        if hasattr(out, "toarray"):
            out = out.toarray()
        return np.asarray(out).reshape(self.mesh.shape())

    def _receivers(self):
        return [self]


class ReceiverSet(MeshRepresentationBase, ReceiverBase):
    def __init__(self, mesh, receivers, **kwargs):
        self.receiver_list = receivers
        ReceiverBase.__init__(self, **kwargs)
        MeshRepresentationBase.__init__(self, mesh, **kwargs)
        self._time_window = ("Set",)
        import scipy.sparse as spsp
        if self._sample_interp_method == "sparse":
            self.sampling_operator = spsp.vstack([r.sampling_operator for r in receivers]).tocsr()
            self.adjoint_sampling_operator = spsp.hstack([r.adjoint_sampling_operator for r in receivers]).tocsc()
        else:
            self.sampling_operator = np.vstack([r.sampling_operator for r in receivers])
            self.adjoint_sampling_operator = np.hstack([r.adjoint_sampling_operator for r in receivers])

    def set_shot(self, shot):
        self.shot = shot
        for r in self.receiver_list:
            r.set_shot(shot)

    @property
    def receiver_count(self):
        return sum(r.receiver_count for r in self.receiver_list)

    def clear_data(self, n):
        self.data = np.zeros((n, self.receiver_count))

    @property
    def data(self):
        return self._data

    @data.setter
    def data(self, value):
        self._data = value
        self.data_version = getattr(self, "data_version", 0) + 1
        c = 0
        if value is not None:
            for r in self.receiver_list:
                r.data = self._data[:, c:c + r.receiver_count]
                c += r.receiver_count

    def sample_data_from_array(self, arr, k=None, data=None):
        v = self.sampling_operator * arr if self._sample_interp_method == "sparse" \
            else np.dot(self.sampling_operator, arr)
        v = np.asarray(v).reshape(1, -1)
        if k is None:
            return v
        if data is None:
            self.data_version = getattr(self, "data_version", 0) + 1
        (self.data if data is None else data)[k, :] = v

    def extend_data_to_array(self, k=None, resid=False, data=None):
        if k is not None:
            if data is None:
                d = self.data[k, :]
            else:
                d = (self.data[k, :] - data[k, :]) if resid else data[k, :]
        else:
            d = data
        d = np.asarray(d).reshape(-1, 1)
        out = self.adjoint_sampling_operator * d if self._sample_interp_method == "sparse" \
            else np.dot(self.adjoint_sampling_operator, d)
        return np.asarray(out).reshape(self.mesh.shape())

    def time_window(self, ts):
        return np.array([r.time_window(ts) for r in self.receiver_list]).T

    def _receivers(self):
        return list(self.receiver_list)


def sampling_entries(receivers, padded_shape):
    """Flat padded indices, receiver ids and weights (times prod(deltas)) of S."""
    return _entries(receivers, padded_shape, scale=True)


def _gathered(receivers):
    """Every receiver's entries in receiver order: (unpadded flat indices,
    receiver ids, weights, prod(deltas) per entry), one mesh conversion for
    all receivers that share the first one's mesh."""
    recs = receivers._receivers()
    idx, chan, w, pd = [], [], [], []
    for r, rec in enumerate(recs):
        i, wt = rec.entries()
        idx.append(np.asarray(i, dtype=np.int64))
        chan.append(np.full(len(idx[-1]), r, dtype=np.int64))
        w.append(wt)
        pd.append(np.full(len(idx[-1]), rec._prod_deltas))
    return recs, np.concatenate(idx), np.concatenate(chan), np.concatenate(w), np.concatenate(pd)


def _positions(recs, idx, chan):
    mesh = recs[0].mesh
    if all(rec.mesh is mesh for rec in recs):
        return padded_positions(mesh, idx)
    return np.concatenate([padded_positions(rec.mesh, idx[chan == r]) for r, rec in enumerate(recs)])


def adjoint_entries(receivers):
    """Padded positions, receiver ids and weights of S* (no prod(deltas))."""
    recs, idx, chan, w, _ = _gathered(receivers)
    return _positions(recs, idx, chan), chan, w


def _entries(receivers, padded_shape, scale):
    recs, idx, chan, w, pd = _gathered(receivers)
    pos = _positions(recs, idx, chan)
    flat = np.ravel_multi_index(tuple(pos.T), tuple(padded_shape))
    # the sampling operator is the weight row times prod(deltas)
    return flat, chan, (w * pd if scale else w)
