"""Shots: a source (or source set) with its receivers (pysit/core/shot.py:12-147),
and source-encoded supershots (pysit/core/shot.py:149-344).

A supershot fires the sources of several shots at once, each scaled by a
random code, and records the same codes' combination of their data; on the
device it is one shot whose injection has one channel per member source.
"""
import copy

import numpy as np

from .sources import PointSource, SourceSet

__all__ = ["Shot", "SourceEncodedSupershot", "EncodedSourceSet"]


class Shot(object):
    def __init__(self, sources, receivers):
        self.receivers = receivers
        receivers.set_shot(self)
        self.sources = sources
        sources.set_shot(self)

    def initialize(self, data_length):
        self.receivers.clear_data(data_length)

    def reset_time_series(self, ts):
        self.sources.reset_time_series(ts)
        self.receivers.reset_time_series(ts)

    def gather(self, as_array=False, offset=None):
        if offset is not None:
            raise NotImplementedError("Gather by offset not yet implemented.")
        return self.receivers.data if as_array else self.receivers


class EncodedSourceSet(SourceSet):
    """A SourceSet whose member sources carry the codes as intensities
    (time coding, shot.py:189-262; frequency coding belongs to the
    frequency-domain solvers, which are not part of this package)."""

    def __init__(self, mesh, sources, time_coded=True, **kwargs):
        if not time_coded:
            raise NotImplementedError("frequency-domain source encoding is not supported (time solvers only)")
        SourceSet.__init__(self, mesh, sources, **kwargs)
        self.time_coded = True
        self.codes = None

    def encode(self, codes, nu=None):
        if nu is not None:
            raise NotImplementedError("frequency-domain source encoding is not supported (time solvers only)")
        codes = np.asarray(codes, dtype=np.float64).reshape(-1)
        if codes.size != len(self.source_list):
            raise ValueError(f"{codes.size} codes for {len(self.source_list)} sources")
        self.codes = codes
        for source, c in zip(self.source_list, codes):
            source.intensity = float(c)


class SourceEncodedSupershot(Shot):
    """Source-encoded supershot of ``shots`` (shot.py:149-344).

    The member shots need PointSource sources, the same receivers (identical
    sampling operators and approximations: a fixed spread) and recorded time
    data.  Each :meth:`encode` draws a new code vector
    (:meth:`generate_weight_vector`: ``'gaussian'`` N(0, 1) or ``'krebs'``
    +-1 with equal probability, from ``rng`` or numpy's global generator as
    the reference does), builds the encoded source set from copies of the
    member sources with the codes as intensities, and the encoded receiver
    set from a copy of the first shot's receivers with data
    sum_i code_i d_i."""

    def __init__(self, shots, weight_type="gaussian", rng=None):
        self.weight_type = weight_type
        self.rng = rng
        self.sequential_shots = shots

    def generate_weight_vector(self):
        n = self._nshots
        if self.weight_type == "gaussian":
            return self.rng.standard_normal(n) if self.rng is not None else np.random.randn(n)
        if self.weight_type == "krebs":
            bits = self.rng.integers(0, 2, n) if self.rng is not None else np.random.randint(0, 2, n)
            return 2 * bits - 1
        raise Exception("Incorrect weight type supplied.")

    def encode(self):
        shots = self._sequential_shots
        sources = []
        for shot in shots:
            src = copy.deepcopy(shot.sources)
            src.set_shot(None)
            sources.append(src)
        enc = EncodedSourceSet(shots[0].sources.mesh, sources, time_coded=True)
        receivers = copy.deepcopy(shots[0].receivers)
        codes = np.asarray(self.generate_weight_vector(), dtype=np.float64)
        data = np.zeros_like(np.asarray(shots[0].receivers.data, dtype=np.float64))
        for shot, c in zip(shots, codes.tolist()):
            data += c * shot.receivers.data
        receivers.data = data
        enc.encode(codes)
        self.codes = codes
        self.sources = enc
        self.receivers = receivers
        enc.set_shot(self)
        receivers.set_shot(self)

    @property
    def sequential_shots(self):
        return self._sequential_shots

    @sequential_shots.setter
    def sequential_shots(self, shots):
        shots = list(shots)
        if not shots:
            raise ValueError("a supershot needs at least one shot")
        for shot in shots:
            if type(shot.sources) is not PointSource:
                raise Exception("The shots have to have a PointSource as source.")
        base = shots[0].receivers
        approx = base.receiver_list[0].approximation
        for shot in shots:
            same = shot.receivers.sampling_operator.shape == base.sampling_operator.shape
            if same:
                d = shot.receivers.sampling_operator - base.sampling_operator
                same = (d.nnz if hasattr(d, "nnz") else np.count_nonzero(d)) == 0
            if not same:
                raise Exception("The receiver acquisition has to be the same for each shot in order to "
                                "construct a supershot.")
            for r in shot.receivers.receiver_list:
                if r.approximation != approx:
                    raise Exception("The receiver approximations are not the same. An encoded stack would "
                                    "make no sense.")
        if shots[0].receivers.data is None:
            raise Exception("No time data recorded yet (frequency data belong to the frequency-domain solvers).")
        self._sequential_shots = shots
        self._fixed_spread_sampling_operator = base.sampling_operator
        self._receiver_approximation = approx
        self._nshots = len(shots)
        self.is_time_simulation = True
        self.encode()
