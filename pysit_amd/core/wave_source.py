"""Source wavelets.

Mirrors pysit/core/wave_source.py:104-203: ``DerivativeGaussianPulse`` is the
n-th derivative of a Gaussian with ``sigma = 1 / (pi nu sqrt(2))``, shifted by
``shift_deviations * sigma``; in time
``w(t) = (-1/sqrt2)^n H_n(x) exp(-x^2)``, ``x = (t - t_shift) / (sqrt2 sigma)``
(physicists' Hermite polynomial), set to zero where ``|w| < threshold``.  The
Ricker wavelet is minus the second derivative.  Values are memoised per
(t, sigma, t_shift, threshold) like the reference's class-level cache.
"""
import math

import numpy as np

__all__ = ["SourceWaveletBase", "DerivativeGaussianPulse", "RickerWavelet", "GaussianPulse"]

_SQRT2 = math.sqrt(2.0)


class SourceWaveletBase(object):
    time_source = False
    frequency_source = False

    def __call__(self, t=None, nu=None, **kwargs):
        if t is not None:
            if self.time_source:
                return self._evaluate_time(t)
            raise TypeError(f"Sources of type {type(self).__name__} are not time-domain sources.")
        if nu is not None:
            if self.frequency_source:
                return self._evaluate_frequency(nu)
            raise TypeError(f"Sources of type {type(self).__name__} are not frequency-domain sources.")
        raise ValueError("Either a time or frequency must be provided.")


class DerivativeGaussianPulse(SourceWaveletBase):
    time_source = True
    frequency_source = True
    _cache = dict()

    def __init__(self, peak_frequency, order=0, threshold=1e-6, shift_deviations=6, t_shift=None):
        self._order = order
        self.peak_frequency = peak_frequency
        self.threshold = threshold
        self.shift_deviations = shift_deviations
        self.sigma = 1 / (math.pi * peak_frequency * _SQRT2)
        self.t_shift = self.shift_deviations * self.sigma if t_shift is None else t_shift
        self._hermite = np.polynomial.Hermite(self.order * [0.0] + [1.0])

    @property
    def order(self):
        return self._order

    def _pulse(self, t):
        key = (t, self.sigma, self.t_shift, self.threshold)
        v = DerivativeGaussianPulse._cache.get(key)
        if v is None:
            x = (t - self.t_shift) / (_SQRT2 * self.sigma)
            v = (-1 / _SQRT2) ** self.order * self._hermite(x) * np.exp(-(x ** 2))
            if np.abs(v) < self.threshold:
                v = 0.0
            DerivativeGaussianPulse._cache[key] = v
        return v

    def _evaluate_time(self, ts):
        if np.iterable(ts):
            return np.array([self._pulse(t) for t in ts])
        return self._pulse(ts)

    def _evaluate_frequency(self, nus):
        scalar = not np.iterable(nus)
        nus = np.atleast_1d(np.asarray(nus, dtype=np.float64))
        omegas = 2 * np.pi * nus
        n = self.order
        v = (math.sqrt(2 * np.pi) * (-1) ** n * (1j * omegas) ** n * self.sigma ** (n + 1)
             * np.exp(-0.5 * (self.sigma ** 2) * omegas ** 2) * np.exp(-1j * 2 * np.pi * nus * self.t_shift))
        v[np.abs(v) < self.threshold] = 0.0
        return v[0] if scalar else v

    _amp_cache = {}

    def amplitudes(self, dt, nsteps):
        """w(k*dt) for k = 0..nsteps-1, each time value formed as k*dt and
        evaluated as a scalar, as the reference's time loop does; the array is
        kept per (wavelet, dt, nsteps), so the shots of a survey evaluate it
        once (a read-only array)."""
        key = (type(self), self.order, self.sigma, self.t_shift, self.threshold, float(dt), int(nsteps))
        a = DerivativeGaussianPulse._amp_cache.get(key)
        if a is None:
            a = np.array([self._evaluate_time(k * dt) for k in range(nsteps)], dtype=np.float64)
            a.setflags(write=False)
            if len(DerivativeGaussianPulse._amp_cache) >= 64:  # time axes change as dt follows the model
                DerivativeGaussianPulse._amp_cache.clear()
            DerivativeGaussianPulse._amp_cache[key] = a
        return a


class RickerWavelet(DerivativeGaussianPulse):
    """Minus the second derivative of a Gaussian, peak frequency ``nu``."""

    def __init__(self, nu, **kwargs):
        kwargs.pop("order", None)
        DerivativeGaussianPulse.__init__(self, nu, order=2, **kwargs)

    def _evaluate_time(self, ts):
        return -1 * DerivativeGaussianPulse._evaluate_time(self, ts)

    def _evaluate_frequency(self, nus):
        return -1 * DerivativeGaussianPulse._evaluate_frequency(self, nus)


class GaussianPulse(DerivativeGaussianPulse):
    def __init__(self, nu, **kwargs):
        kwargs.pop("order", None)
        DerivativeGaussianPulse.__init__(self, nu, order=0, **kwargs)
