"""Objective functions."""
from .temporal_least_squares import *  # noqa: F401,F403
