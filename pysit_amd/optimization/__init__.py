"""Optimization loops."""
from .optimization import *  # noqa: F401,F403
from .lbfgs import *  # noqa: F401,F403
