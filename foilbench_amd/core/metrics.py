"""Spectrum of the benchmark runner's wake probe (fb/core/metrics.py:14-21, 107-150): a Hann-windowed
periodogram of the transverse velocity samples (one value per output step), on the host.

The field diagnostics of the same reference module (energy, enstrophy, divergence, wake width, recirculation
area) are computed on the device by each solver's diagnostics().
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True, slots=True)
class WakeSpectrum:
    sample_count: int
    frequency_resolution: float
    transverse_rms: float
    dominant_frequency: float
    strouhal_number: float
    dominant_power_fraction: float


def analyze_wake_probe(v: np.ndarray, sample_dt: float, chord: float, freestream_speed: float) -> WakeSpectrum:
    if v.ndim != 1:
        raise ValueError("wake probe samples must be one-dimensional")
    if v.size < 8:
        raise ValueError("wake spectrum requires at least eight samples")
    if sample_dt <= 0.0 or chord <= 0.0 or freestream_speed <= 0.0:
        raise ValueError("sample_dt, chord, and freestream_speed must be positive")
    if not np.isfinite(v).all():
        raise ValueError("wake probe samples must be finite")
    c = v - np.mean(v)
    rms = float(np.sqrt(np.mean(c * c)))
    power = np.abs(np.fft.rfft(c * np.hanning(c.size))) ** 2
    power[0] = 0.0
    total = float(np.sum(power))
    res = 1.0 / (c.size * sample_dt)
    if total <= np.finfo(np.float64).tiny:
        return WakeSpectrum(v.size, res, rms, 0.0, 0.0, 0.0)
    k = int(np.argmax(power))
    f = float(np.fft.rfftfreq(c.size, sample_dt)[k])
    return WakeSpectrum(v.size, res, rms, f, f * chord / freestream_speed, float(power[k] / total))
