"""Solvers: registration factory, model parameters, the CDA time solvers."""
from .factory import *  # noqa: F401,F403
from .model_parameter import *  # noqa: F401,F403
from .solver_base import *  # noqa: F401,F403
from .cda_time import *  # noqa: F401,F403
