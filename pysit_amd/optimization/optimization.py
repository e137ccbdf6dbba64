"""Descent-loop driver.

Mirrors pysit/optimization/optimization.py:16-501: histories (value,
gradient, objective, step, alpha, run time...), the outer ``__call__`` over
iteration counts or (steps, objective_arguments) pairs, ``inner_loop``
(gradient with objective value, step selection, model update in linear
space) and the backtracking line search: alpha is halved until the trial
model validates (finite, positive), accepted when
f(x + a d) <= f(x) + c a <g, a d> (c = 1e-3; note the step enters the inner
product already scaled) or within a relative 1e-6, else shrunk by 0.8, at
most ``max_linesearch_iterations`` passes.
"""
import copy
import sys
import time

import numpy as np

__all__ = ["OptimizationBase"]


class OptimizationBase(object):
    def __init__(self, objective):
        self.objective_function = objective
        self.solver = objective.solver
        self.verbose = False
        self.use_parallel = objective.use_parallel()
        self.max_linesearch_iterations = 10
        self.logfile = sys.stdout

    def reset(self, append_mode, value_frequency=0, gradient_frequency=0, gradient_length_frequency=0,
              step_frequency=0, step_length_frequency=0, residual_length_frequency=0,
              objective_frequency=0, run_time_frequency=0, alpha_frequency=1, *args, **kwargs):
        if not append_mode or not hasattr(self, "iteration"):
            self.base_model = self.solver.ModelParameters(self.solver.mesh)
            self.iteration = 0
            for name, f in (("value", value_frequency), ("gradient", gradient_frequency),
                            ("gradient_length", gradient_length_frequency), ("step", step_frequency),
                            ("step_length", step_length_frequency),
                            ("residual_length", residual_length_frequency),
                            ("objective", objective_frequency), ("run_time", run_time_frequency),
                            ("alpha", alpha_frequency)):
                self.init_history(name, f)

    def init_history(self, arg, freq):
        setattr(self, arg + "_history", {})
        setattr(self, arg + "_frequency", freq)

    def query_store_history(self, arg, i):
        f = getattr(self, arg + "_frequency")
        return f and (not np.mod(i, f))

    def store_history(self, arg, i, val, force=False):
        if self.use_parallel and self.objective_function.parallel_wrap_shot.rank != 0:
            return
        f = getattr(self, arg + "_frequency")
        if f and (force or not np.mod(i, f)):
            getattr(self, arg + "_history")[i] = copy.deepcopy(val)

    def retrieve_history(self, arg):
        if getattr(self, arg + "_frequency"):
            return list(zip(*sorted(getattr(self, arg + "_history").items())))
        return None, None

    def _print(self, *args):
        if self.use_parallel and self.objective_function.parallel_wrap_shot.rank != 0:
            return
        if self.verbose:
            print(*args, file=self.logfile)

    def initialize(self, initial_value, **kwargs):
        self.base_model = copy.deepcopy(initial_value)
        self.solver.model_parameters = self.base_model

    def __call__(self, shots, initial_value, iteration_parameters, line_search="backtrack", verbose=False,
                 append=False, status_configuration={}, **kwargs):
        self.reset(append, **status_configuration)
        self.verbose = verbose
        self.line_search = line_search
        if isinstance(line_search, str):
            self.ls_method, self.ls_config = line_search, None
        else:
            self.ls_method, self.ls_config = line_search[0], line_search[1:]
        self.initialize(initial_value, **kwargs)
        if np.iterable(iteration_parameters):
            for ip in iteration_parameters:
                if isinstance(ip, (tuple, list)):
                    steps, arguments = ip
                else:
                    raise ValueError(f"Invalid iteration parameter {ip} detected.")
                self.inner_loop(shots, steps, objective_arguments=arguments, **kwargs)
        elif isinstance(iteration_parameters, int):
            self.inner_loop(shots, iteration_parameters, **kwargs)
        else:
            raise ValueError(f"Singular iteration parameters of type {type(iteration_parameters)} are not permitted.")
        return self.base_model

    def inner_loop(self, shots, steps, objective_arguments={}, **kwargs):
        for _ in range(steps):
            tt = time.time()
            i = self.iteration
            self.store_history("value", i, self.base_model)
            self._print(f"Iteration {i}")
            self.solver.model_parameters = self.base_model
            aux_info = {"objective_value": (True, None), "residual_norm": (True, None)}
            objective_arguments.update(kwargs)
            gradient = self.objective_function.compute_gradient(shots, self.base_model, aux_info=aux_info,
                                                                **objective_arguments)
            objective_value = aux_info["objective_value"][1]
            self.store_history("gradient", i, gradient)
            gnorm = gradient.norm()
            self._print(f"  gradnorm {gnorm}")
            self.store_history("gradient_length", i, gnorm)
            if objective_value is not None:
                self.store_history("objective", i, objective_value)
                self._print(f"  objective {objective_value}")
            if aux_info["residual_norm"][1] is not None:
                self.store_history("residual_length", i, aux_info["residual_norm"][1])
            step = self._select_step(shots, objective_value, gradient, i, objective_arguments, **kwargs)
            self.store_history("step_length", i, step.norm())
            self.store_history("step", i, step)
            self.base_model += step
            self.store_history("run_time", i, time.time() - tt)
            self.iteration += 1
            self._print(f"  run time {time.time() - tt}s")

    def _select_step(self, shots, current_objective_value, gradient, iteration, objective_arguments, **kwargs):
        raise NotImplementedError("_select_step must be implemented by a subclass.")

    def select_alpha(self, shots, gradient, direction, objective_arguments, **kwargs):
        if self.ls_method == "constant":
            return self.ls_config[0]
        if self.ls_method == "backtrack":
            return self._backtrack_line_search(shots, gradient, direction, objective_arguments, **kwargs)
        if self.ls_method == "linear":
            raise NotImplementedError("Linear selection of alpha is an objective function dependent operation.")
        raise ValueError(f"Alpha selection method {self.ls_method} invalid")

    def _backtrack_line_search(self, shots, gradient, direction, objective_arguments,
                               current_objective_value=None, alpha0_kwargs={}, **kwargs):
        shrink = 0.8
        c = 1e-3
        fp_rel = 1e-6
        if current_objective_value is None:
            fk = self.objective_function.evaluate(shots, self.base_model, **objective_arguments)
        else:
            fk = current_objective_value
        a0_kwargs = dict(alpha0_kwargs)
        a0_kwargs["upscale_factor"] = 0.7
        alpha = self._compute_alpha0(current_objective_value, gradient, **a0_kwargs)
        passes = 1
        self._print("  Starting: ", alpha, fk)
        while True:
            alpha *= 2
            valid = False
            while not valid:
                alpha /= 2
                tdir = alpha * direction
                model = self.base_model + tdir
                valid = model.validate()
            self.solver.model_parameters = model
            fkp1 = self.objective_function.evaluate(shots, model, **objective_arguments)
            cmp = fk + alpha * c * gradient.inner_product(tdir)
            self._print(f"  Pass {passes}: a:{alpha}; {fkp1} ?<= {cmp}")
            if fkp1 <= cmp or abs(fkp1 - cmp) / abs(fkp1) <= fp_rel:
                break
            if passes > self.max_linesearch_iterations:
                self._print(f"Too many passes ({passes}), attempting to use current alpha ({alpha}).")
                break
            passes += 1
            alpha = alpha * shrink
        self.prev_alpha = alpha
        return alpha
