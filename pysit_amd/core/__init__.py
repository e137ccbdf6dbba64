"""Geometry, acquisition and wavelets (the host side of pysit.core)."""
from .domain import *  # noqa: F401,F403
from .mesh import *  # noqa: F401,F403
from .point_ops import *  # noqa: F401,F403
from .sources import *  # noqa: F401,F403
from .receivers import *  # noqa: F401,F403
from .wave_source import *  # noqa: F401,F403
from .shot import *  # noqa: F401,F403
from .acquisition import *  # noqa: F401,F403
