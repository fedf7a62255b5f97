"""NACA four-digit foil: scalar properties on the host, distance queries on the GPU.

Mirrors ``NacaFoil`` (fb/core/geometry.py:12-159). Scalars (camber, thickness, maximum
radius, rotation cos/sin, wall velocity of given points) are evaluated by the host in the
reference's expression order; the per-point and per-cell signed distance and normals run in
the gfx950 kernels of csrc/geometry.hip.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .. import _native as N
from .models import ControlState, DomainSpec, FoilSpec


def foil_struct(spec: FoilSpec) -> N.fb_foil:
    """Scalar NACA coefficients, evaluated as Python floats like fb/core/geometry.py:60-90."""
    m = int(spec.naca[0]) / 100.0
    p = int(spec.naca[1]) / 10.0
    t = int(spec.naca[2:]) / 100.0
    chord = float(spec.chord)
    f = N.fb_foil()
    f.chord = chord
    f.pivot_x = float(spec.pivot[0])
    f.pivot_y = float(spec.pivot[1])
    f.half_scale = 5.0 * t * chord
    cambered = m > 0.0 and p > 0.0
    f.cambered = 1 if cambered else 0
    f.camber_fwd = m / p**2 if cambered else 0.0
    f.camber_aft = m / (1.0 - p) ** 2 if cambered else 0.0
    f.two_p = 2.0 * p
    f.one_minus_two_p = 1.0 - 2.0 * p
    f.p = p
    return f


def pose_struct(angle_degrees: float) -> N.fb_pose:
    """cos/sin of np.deg2rad(angle) as fb/core/geometry.py:36-38 evaluates them."""
    a = np.deg2rad(angle_degrees)
    ps = N.fb_pose()
    ps.angle_degrees = float(angle_degrees)
    ps.cos_a = float(np.cos(a))
    ps.sin_a = float(np.sin(a))
    return ps


def domain_struct(domain: DomainSpec) -> N.fb_domain:
    d = N.fb_domain()
    d.nx = domain.nx
    d.ny = domain.ny
    d.periodic_x = 1 if "x" in domain.periodic_axes else 0
    d.periodic_y = 1 if "y" in domain.periodic_axes else 0
    d.x0, d.x1 = (float(v) for v in domain.bounds[0])
    d.y0, d.y1 = (float(v) for v in domain.bounds[1])
    return d


@dataclass(frozen=True, slots=True)
class NacaFoil:
    spec: FoilSpec

    @property
    def maximum_camber(self) -> float:
        return int(self.spec.naca[0]) / 100.0

    @property
    def camber_position(self) -> float:
        return int(self.spec.naca[1]) / 10.0

    @property
    def thickness(self) -> float:
        return int(self.spec.naca[2:]) / 100.0

    @property
    def maximum_radius(self) -> float:
        # fb/core/geometry.py:28-33
        vertical = (self.maximum_camber + 0.51 * self.thickness) * self.spec.chord
        return float(np.hypot(0.75 * self.spec.chord, vertical))

    def signed_distance(self, points: np.ndarray, angle_degrees: float, ctx: N.Context | None = None) -> np.ndarray:
        ctx = ctx or N.Context.default()
        pts = np.ascontiguousarray(points[:, :2], dtype=np.float64)
        out = np.empty(pts.shape[0], dtype=np.float64)
        f = foil_struct(self.spec)
        ps = pose_struct(angle_degrees)
        N.check(ctx.lib.fb_naca_points(ctx.handle, C.byref(f), C.byref(ps), pts.shape[0], N.ptr(pts), N.ptr(out), None))
        return out.astype(points.dtype, copy=False)

    def normals(self, points: np.ndarray, angle_degrees: float, ctx: N.Context | None = None) -> np.ndarray:
        ctx = ctx or N.Context.default()
        pts = np.ascontiguousarray(points[:, :2], dtype=np.float64)
        sdf = np.empty(pts.shape[0], dtype=np.float64)
        nrm = np.empty((pts.shape[0], 2), dtype=np.float64)
        f = foil_struct(self.spec)
        ps = pose_struct(angle_degrees)
        N.check(ctx.lib.fb_naca_points(ctx.handle, C.byref(f), C.byref(ps), pts.shape[0], N.ptr(pts), N.ptr(sdf),
                                       N.ptr(nrm)))
        return nrm.astype(points.dtype, copy=False)

    def contains(self, points: np.ndarray, angle_degrees: float, ctx: N.Context | None = None) -> np.ndarray:
        return self.signed_distance(points, angle_degrees, ctx) <= 0.0

    def wall_velocity(self, points: np.ndarray, control: ControlState) -> np.ndarray:
        # fb/core/geometry.py:131-139
        rel = points[:, :2] - np.asarray(self.spec.pivot[:2], dtype=points.dtype)
        w = np.deg2rad(control.angular_velocity_degrees)
        out = np.empty_like(rel)
        out[:, 0] = -w * rel[:, 1]
        out[:, 1] = w * rel[:, 0]
        return out

    def cell_sdf(self, domain: DomainSpec, angle_degrees: float, ctx: N.Context | None = None):
        """(signed distance, solid mask) of the cell centres (fb/core/geometry.py:157-168)."""
        ctx = ctx or N.Context.default()
        sdf = np.empty((domain.ny, domain.nx), dtype=np.float64)
        solid = np.empty((domain.ny, domain.nx), dtype=np.uint8)
        d = domain_struct(domain)
        f = foil_struct(self.spec)
        ps = pose_struct(angle_degrees)
        N.check(ctx.lib.fb_naca_cell_sdf(ctx.handle, C.byref(d), C.byref(f), C.byref(ps), N.ptr(sdf), N.ptr(solid)))
        return sdf, solid.astype(bool)

    def mask(self, domain: DomainSpec, angle_degrees: float, ctx: N.Context | None = None) -> np.ndarray:
        return self.cell_sdf(domain, angle_degrees, ctx)[1]


def cell_centers(domain: DomainSpec) -> np.ndarray:
    """Cell-centre coordinates (ny, nx, 2) as fb/core/geometry.py:162-168 builds them."""
    x0, x1 = domain.bounds[0]
    y0, y1 = domain.bounds[1]
    x = np.linspace(x0 + 0.5 * domain.dx, x1 - 0.5 * domain.dx, domain.nx)
    y = np.linspace(y0 + 0.5 * domain.dy, y1 - 0.5 * domain.dy, domain.ny)
    xx, yy = np.meshgrid(x, y)
    return np.stack((xx, yy), axis=-1)
