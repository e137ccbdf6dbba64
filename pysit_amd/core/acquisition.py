"""Equally spaced acquisition geometry (pysit/core/acquisition.py:14-103).

Receivers: ``linspace(xmin, xmax, receivers)`` (2D) or an x-y grid (3D) at
``receiver_depth``; sources at ``xmin + (xmax - xmin) * (i + 1) / (sources + 1)``
at ``source_depth``; each shot gets its own copy of the receiver set.
"""
import copy

import numpy as np

from .receivers import PointReceiver, ReceiverSet
from .shot import Shot
from .sources import PointSource

__all__ = ["equispaced_acquisition"]


class _SerialShots(object):
    size = 1
    rank = 0


def equispaced_acquisition(mesh, wavelet, sources=1, receivers="max", source_depth=None,
                           source_kwargs={}, receiver_depth=None, receiver_kwargs={},
                           parallel_shot_wrap=None):
    wrap = parallel_shot_wrap or _SerialShots()
    d = mesh.domain
    xmin, xmax = d.x.lbound, d.x.rbound
    zmin = d.z.lbound
    source_depth = zmin if source_depth is None else source_depth
    receiver_depth = zmin if receiver_depth is None else receiver_depth
    shots = []
    if mesh.dim == 2:
        receivers = mesh.x.n if receivers == "max" else receivers
        sources = mesh.x.n if sources == "max" else sources
        if receivers > mesh.x.n:
            raise ValueError("Number of receivers exceeds mesh nodes.")
        if sources > mesh.x.n:
            raise ValueError("Number of sources exceeds mesh nodes.")
        base = ReceiverSet(mesh, [PointReceiver(mesh, (x, receiver_depth), **receiver_kwargs)
                                  for x in np.linspace(xmin, xmax, receivers)])
        local = (sources // wrap.size, 1)
    elif mesh.dim == 3:
        ymin, ymax = d.y.lbound, d.y.rbound
        receivers = (mesh.x.n, mesh.y.n) if receivers == "max" else receivers
        sources = (mesh.x.n, mesh.y.n) if sources == "max" else sources
        if receivers[0] > mesh.x.n or receivers[1] > mesh.y.n:
            raise ValueError("Number of receivers exceeds mesh nodes.")
        if sources[0] > mesh.x.n or sources[1] > mesh.y.n:
            raise ValueError("Number of sources exceeds mesh nodes.")
        base = ReceiverSet(mesh, [PointReceiver(mesh, (x, y, receiver_depth), **receiver_kwargs)
                                  for x in np.linspace(xmin, xmax, receivers[0])
                                  for y in np.linspace(ymin, ymax, receivers[1])])
        local = (sources[0] // wrap.size, sources[1] // wrap.size)
    else:
        raise ValueError("equispaced_acquisition needs a 2D or 3D mesh")
    for i in range(int(local[0])):
        for j in range(int(local[1])):
            ii = i + local[0] * wrap.rank
            jj = j + local[1] * wrap.rank
            if mesh.dim == 2:
                pos = (xmin + (xmax - xmin) * (ii + 1.0) / (sources + 1.0), source_depth)
            else:
                pos = (xmin + (xmax - xmin) * (ii + 1.0) / (sources[0] + 1.0),
                       d.y.lbound + (d.y.rbound - d.y.lbound) * (jj + 1.0) / (sources[1] + 1.0),
                       source_depth)
            shots.append(Shot(PointSource(mesh, pos, wavelet, **source_kwargs), copy.deepcopy(base)))
    return shots
