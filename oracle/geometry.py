"""NACA four-digit geometry restated from fb/core/geometry.py (test infrastructure only).

Pinned by: spec/conformance/naca2412.json (surfaces, signed distance, normals, contains at
3 angles) and spec/conformance/geometry-v1.json (NACA 0012, chord 2.5, 23 degrees, wall
velocity, maximum radius), copied under tests/golden/.
"""

from __future__ import annotations

import numpy as np


class NacaFoil:
    """Restates ``NacaFoil`` (fb/core/geometry.py:12-159)."""

    def __init__(self, naca: str, chord: float, pivot: tuple[float, float]) -> None:
        self.naca = naca
        self.chord = float(chord)
        self.pivot = (float(pivot[0]), float(pivot[1]))

    # fb/core/geometry.py:16-26
    @property
    def maximum_camber(self) -> float:
        return int(self.naca[0]) / 100.0

    @property
    def camber_position(self) -> float:
        return int(self.naca[1]) / 10.0

    @property
    def thickness(self) -> float:
        return int(self.naca[2:]) / 100.0

    # fb/core/geometry.py:28-33
    @property
    def maximum_radius(self) -> float:
        vertical = (self.maximum_camber + 0.51 * self.thickness) * self.chord
        return float(np.hypot(0.75 * self.chord, vertical))

    # fb/core/geometry.py:35-44
    def to_local(self, points: np.ndarray, angle_degrees: float) -> np.ndarray:
        angle = np.deg2rad(angle_degrees)
        c = np.cos(angle)
        s = np.sin(angle)
        t = points[:, :2] - np.asarray(self.pivot, dtype=points.dtype)
        local = np.empty_like(t)
        local[:, 0] = c * t[:, 0] + s * t[:, 1]
        local[:, 1] = -s * t[:, 0] + c * t[:, 1]
        local[:, 0] += 0.25 * self.chord
        return local

    # fb/core/geometry.py:60-90
    def surfaces(self, x_local: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        chord = self.chord
        x = np.clip(x_local / chord, 0.0, 1.0)
        half = 5.0 * self.thickness * chord * (
            0.2969 * np.sqrt(np.maximum(x, 0.0))
            - 0.1260 * x
            - 0.3516 * x**2
            + 0.2843 * x**3
            - 0.1036 * x**4
        )
        m = self.maximum_camber
        p = self.camber_position
        camber = np.zeros_like(x)
        if m > 0.0 and p > 0.0:
            fwd = x < p
            camber[fwd] = m / p**2 * (2.0 * p * x[fwd] - x[fwd] ** 2)
            camber[~fwd] = m / (1.0 - p) ** 2 * ((1.0 - 2.0 * p) + 2.0 * p * x[~fwd] - x[~fwd] ** 2)
            camber *= chord
        return camber + half, camber - half

    # fb/core/geometry.py:92-106
    def signed_distance(self, points: np.ndarray, angle_degrees: float) -> np.ndarray:
        local = self.to_local(points, angle_degrees)
        upper, lower = self.surfaces(local[:, 0])
        x = local[:, 0]
        y = local[:, 1]
        inside_x = (x >= 0.0) & (x <= self.chord)
        outside_v = np.maximum(y - upper, lower - y)
        inside_v = -np.minimum(upper - y, y - lower)
        vertical = np.where((y <= upper) & (y >= lower), inside_v, outside_v)
        dx = np.maximum(-x, x - self.chord)
        outside_x = np.maximum(dx, 0.0)
        d = np.where(inside_x, vertical, np.hypot(outside_x, np.maximum(vertical, 0.0)))
        return np.asarray(d, dtype=points.dtype)

    # fb/core/geometry.py:108-126
    def normals(self, points: np.ndarray, angle_degrees: float) -> np.ndarray:
        eps = max(self.chord * 1.0e-4, 1.0e-6)
        ox = np.asarray([eps, 0.0], dtype=points.dtype)
        oy = np.asarray([0.0, eps], dtype=points.dtype)
        gx = self.signed_distance(points + ox, angle_degrees) - self.signed_distance(points - ox, angle_degrees)
        gy = self.signed_distance(points + oy, angle_degrees) - self.signed_distance(points - oy, angle_degrees)
        g = np.stack((gx, gy), axis=1)
        n = np.linalg.norm(g, axis=1, keepdims=True)
        amb = n[:, 0] < eps
        if np.any(amb):
            a = np.deg2rad(angle_degrees)
            g[amb, 0] = -np.sin(a)
            g[amb, 1] = np.cos(a)
            n[amb, 0] = 1.0
        return g / np.maximum(n, eps)

    def contains(self, points: np.ndarray, angle_degrees: float) -> np.ndarray:
        return self.signed_distance(points, angle_degrees) <= 0.0

    # fb/core/geometry.py:131-139
    def wall_velocity(self, points: np.ndarray, angular_velocity_degrees: float) -> np.ndarray:
        r = points[:, :2] - np.asarray(self.pivot, dtype=points.dtype)
        w = np.deg2rad(angular_velocity_degrees)
        out = np.empty_like(r)
        out[:, 0] = -w * r[:, 1]
        out[:, 1] = w * r[:, 0]
        return out

    # fb/core/geometry.py:157-159
    def mask(self, bounds, nx: int, ny: int, angle_degrees: float) -> np.ndarray:
        pts = cell_centers(bounds, nx, ny)
        return self.contains(pts.reshape(-1, 2), angle_degrees).reshape(ny, nx)


def cell_centers(bounds, nx: int, ny: int) -> np.ndarray:
    """fb/core/geometry.py:162-168 (np.linspace of the centre coordinates)."""
    (x0, x1), (y0, y1) = bounds
    dx = (x1 - x0) / nx
    dy = (y1 - y0) / ny
    x = np.linspace(x0 + 0.5 * dx, x1 - 0.5 * dx, nx)
    y = np.linspace(y0 + 0.5 * dy, y1 - 0.5 * dy, ny)
    xx, yy = np.meshgrid(x, y)
    return np.stack((xx, yy), axis=-1)
