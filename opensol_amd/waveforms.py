"""Temporal envelopes of the soft sources (core/sources.py:23-100 of the reference): a Python float time gives a
Python float (math path), an array of times a float64 array."""
import math

import numpy as np


class GaussianPulse:
    """exp(-((t - t0)/tau)^2) sin(2 pi f0 (t - t0)), tau = 1/(bandwidth f0 pi), t0 = t0_factor tau."""

    def __init__(self, f0: float, bandwidth: float = 0.5, t0_factor: float = 4.5):
        self.f0 = float(f0)
        self.bandwidth = float(bandwidth)
        self.tau = 1.0 / (self.bandwidth * self.f0 * np.pi)
        self.t0 = t0_factor * self.tau

    def __call__(self, t):
        if isinstance(t, (float, int)):
            delay = t - self.t0
            return math.exp(-(delay / self.tau) ** 2) * math.sin(2.0 * math.pi * self.f0 * delay)
        delay = np.asarray(t, dtype=np.float64) - self.t0
        return np.exp(-(delay / self.tau) ** 2) * np.sin(2.0 * np.pi * self.f0 * delay)

    @property
    def duration(self) -> float:
        return 2.0 * self.t0

    def spectral_limit(self) -> float:
        """Frequency above which the pulse spectrum stays below 1e-6 of its peak."""
        return self.f0 * (1.0 + 3.72 * self.bandwidth)


class ContinuousWave:
    """(1 - exp(-t/tau_ramp)) sin(2 pi f0 t), tau_ramp = ramp_periods / f0."""

    def __init__(self, f0: float, ramp_periods: float = 3.0):
        self.f0 = float(f0)
        self.tau_ramp = ramp_periods / self.f0

    def __call__(self, t):
        if isinstance(t, (float, int)):
            return (1.0 - math.exp(-t / self.tau_ramp)) * math.sin(2.0 * math.pi * self.f0 * t)
        times = np.asarray(t, dtype=np.float64)
        return (1.0 - np.exp(-times / self.tau_ramp)) * np.sin(2.0 * np.pi * self.f0 * times)

    def spectral_limit(self):
        return None


class RickerWavelet:
    """(1 - 2 a^2) exp(-a^2), a = pi f_peak (t - t0), t0 = t0_factor / f_peak."""

    def __init__(self, f_peak: float, t0_factor: float = 4.0):
        self.f_peak = float(f_peak)
        self.t0 = t0_factor / self.f_peak

    @property
    def f0(self) -> float:
        return self.f_peak

    def __call__(self, t):
        if isinstance(t, (float, int)):
            arg = math.pi * self.f_peak * (t - self.t0)
            return (1.0 - 2.0 * arg * arg) * math.exp(-arg * arg)
        arg = np.pi * self.f_peak * (np.asarray(t, dtype=np.float64) - self.t0)
        return (1.0 - 2.0 * arg ** 2) * np.exp(-arg ** 2)

    @property
    def duration(self) -> float:
        return 2.0 * self.t0

    def spectral_limit(self) -> float:
        return 4.2 * self.f_peak


def make_temporal(waveform, frequency, bandwidth):
    """physics/em/__init__.py:580-587."""
    if waveform == 'gaussian':
        return GaussianPulse(frequency, bandwidth)
    if waveform == 'cw':
        return ContinuousWave(frequency)
    if waveform == 'ricker':
        return RickerWavelet(frequency)
    raise ValueError(f"unknown waveform {waveform!r}; valid: 'gaussian', 'cw', 'ricker'")
