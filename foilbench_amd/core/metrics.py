"""Run-level diagnostics on exported fields (fb/core/metrics.py:24-150).

These run once per benchmark run on host copies of the exported fields, like the
reference's; they are not part of the per-step update.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .models import DomainSpec


@dataclass(frozen=True, slots=True)
class WakeSpectrum:
    sample_count: int
    frequency_resolution: float
    transverse_rms: float
    dominant_frequency: float
    strouhal_number: float
    dominant_power_fraction: float


def speed_squared(velocity: np.ndarray) -> np.ndarray:
    return np.sum(velocity * velocity, axis=-1)


def kinetic_energy(velocity: np.ndarray) -> float:
    return 0.5 * float(np.mean(speed_squared(velocity)))


def vorticity(velocity: np.ndarray, domain: DomainSpec) -> np.ndarray:
    du_dy = np.gradient(velocity[:, :, 0], domain.dy, axis=0)
    dv_dx = np.gradient(velocity[:, :, 1], domain.dx, axis=1)
    return dv_dx - du_dy


def enstrophy(velocity: np.ndarray, domain: DomainSpec) -> float:
    w = vorticity(velocity, domain)
    return 0.5 * float(np.mean(w * w))


def divergence_l2(velocity: np.ndarray, domain: DomainSpec) -> float:
    div = np.gradient(velocity[:, :, 0], domain.dx, axis=1) + np.gradient(velocity[:, :, 1], domain.dy, axis=0)
    return float(np.sqrt(np.mean(div * div)))


def _centers_x(domain: DomainSpec) -> np.ndarray:
    return np.linspace(domain.bounds[0][0] + 0.5 * domain.dx, domain.bounds[0][1] - 0.5 * domain.dx, domain.nx)


def wake_width(velocity, domain, pivot_x, chord=1.0, freestream_u=1.0, solid=None, threshold=0.1) -> float:
    cols = _centers_x(domain) > pivot_x + chord
    if not np.any(cols):
        return 0.0
    active = freestream_u - velocity[:, cols, 0] > threshold * abs(freestream_u)
    if solid is not None:
        active &= ~solid[:, cols]
    return float(np.count_nonzero(np.any(active, axis=1)) * domain.dy)


def recirculation_area(velocity, domain, pivot_x, solid=None) -> float:
    down = _centers_x(domain) > pivot_x
    rec = velocity[:, down, 0] < 0.0
    if solid is not None:
        rec &= ~solid[:, down]
    return float(np.count_nonzero(rec) * domain.dx * domain.dy)


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


def bilinear_sample(field: np.ndarray, points: np.ndarray, domain: DomainSpec, offset=(0.5, 0.5)) -> np.ndarray:
    """sample_staggered_scalar (fb/core/interpolation.py:9-52) for a host field copy."""
    x0 = domain.bounds[0][0]
    y0 = domain.bounds[1][0]
    gx = (points[:, 0] - x0) / domain.dx - offset[0]
    gy = (points[:, 1] - y0) / domain.dy - offset[1]
    px = "x" in domain.periodic_axes
    py = "y" in domain.periodic_axes
    gx = np.mod(gx, domain.nx) if px else np.clip(gx, 0.0, field.shape[1] - 1.0)
    gy = np.mod(gy, domain.ny) if py else np.clip(gy, 0.0, field.shape[0] - 1.0)
    ix0 = np.floor(gx).astype(np.int64)
    iy0 = np.floor(gy).astype(np.int64)
    tx = gx - ix0
    ty = gy - iy0
    ix1 = np.mod(ix0 + 1, domain.nx) if px else np.minimum(ix0 + 1, field.shape[1] - 1)
    iy1 = np.mod(iy0 + 1, domain.ny) if py else np.minimum(iy0 + 1, field.shape[0] - 1)
    out = ((1.0 - tx) * (1.0 - ty) * field[iy0, ix0] + tx * (1.0 - ty) * field[iy0, ix1]
           + (1.0 - tx) * ty * field[iy1, ix0] + tx * ty * field[iy1, ix1])
    return np.asarray(out, dtype=field.dtype)


def sample_vector(field: np.ndarray, points: np.ndarray, domain: DomainSpec) -> np.ndarray:
    return np.stack([bilinear_sample(field[:, :, k], points, domain) for k in range(field.shape[2])], axis=1)
