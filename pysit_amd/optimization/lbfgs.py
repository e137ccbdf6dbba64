"""Limited-memory BFGS (pysit/optimization/lbfgs.py:15-124).

Two-loop recursion with the mesh inner product; the memory pairs
(rho, s, y) are completed lazily at the next iteration (y = g_{k+1} - g_k,
s = x_{k+1} - x_k taken from the models, since the model update may be
post-processed); gamma = <s, y> / <y, y>.  The first step of an inner loop
uses alpha0 = f / (||g|| * prod(deltas))^2, later steps alpha0 = 1.
"""
import copy
from collections import deque

import numpy as np

from .optimization import OptimizationBase

__all__ = ["LBFGS"]


class LBFGS(OptimizationBase):
    def __init__(self, objective, memory_length=None, reset_on_new_inner_loop_call=True, *args, **kwargs):
        OptimizationBase.__init__(self, objective, *args, **kwargs)
        self.prev_alpha = None
        self.prev_model = None
        self.memory_length = memory_length
        self.reset_on_new_inner_loop_call = reset_on_new_inner_loop_call
        self._reset_memory()

    def _reset_memory(self):
        self.memory = deque([], maxlen=self.memory_length)
        self._reset_line_search = True
        self.prev_model = None

    def inner_loop(self, *args, **kwargs):
        if self.reset_on_new_inner_loop_call:
            self._reset_memory()
        OptimizationBase.inner_loop(self, *args, **kwargs)

    def _select_step(self, shots, current_objective_value, gradient, iteration, objective_arguments, **kwargs):
        mem = self.memory
        q = copy.deepcopy(gradient)
        x_k = copy.deepcopy(self.base_model)
        if len(mem) > 0:
            mem[-1][2] += gradient
            mem[-1][1] = x_k - self.prev_model
            mem[-1][0] = 1.0 / mem[-1][2].inner_product(mem[-1][1])
            gamma = mem[-1][1].inner_product(mem[-1][2]) / mem[-1][2].inner_product(mem[-1][2])
        else:
            gamma = 1.0
        alphas = []
        for rho, s, y in reversed(mem):
            a = rho * s.inner_product(q)
            q -= a * y
            alphas.append(a)
        alphas.reverse()
        r = gamma * q
        for a, (rho, s, y) in zip(alphas, mem):
            beta = rho * y.inner_product(r)
            r += (a - beta) * s
        direction = -1.0 * r
        alpha0_kwargs = {"reset": False}
        if self._reset_line_search:
            alpha0_kwargs = {"reset": True}
            self._reset_line_search = False
        alpha = self.select_alpha(shots, gradient, direction, objective_arguments,
                                  current_objective_value=current_objective_value,
                                  alpha0_kwargs=alpha0_kwargs, **kwargs)
        self._print(f"  alpha {alpha}")
        self.store_history("alpha", iteration, alpha)
        step = alpha * direction
        self.prev_model = x_k
        self.memory.append([None, None, copy.deepcopy(-1 * gradient)])
        return step

    def _compute_alpha0(self, phi0, grad0, reset=False, *args, **kwargs):
        if reset:
            return phi0 / (grad0.norm() * np.prod(self.solver.mesh.deltas)) ** 2
        return 1.0
