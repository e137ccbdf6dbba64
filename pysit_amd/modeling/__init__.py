"""Time-domain modeling (forward, adjoint, migration, Born) on the GPU."""
from .temporal_modeling import *  # noqa: F401,F403
from .data_modeling import *  # noqa: F401,F403
