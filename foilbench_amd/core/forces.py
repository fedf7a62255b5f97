"""Lift and drag from a force vector on the foil (build-defined: the reference computes none).

Drag is the component along the freestream direction, lift the component along the freestream
rotated by +90 degrees; coefficients use the dynamic pressure 0.5 rho U^2 and the chord (per unit
depth). The solvers' `forces()` methods obtain the force vector on the device in FP64
(`fb_mac_forces`, `fb_lbm_forces`).
"""

from __future__ import annotations

import math


def lift_drag(fx: float, fy: float, freestream: tuple[float, float], chord: float, rho: float = 1.0,
              speed: float | None = None) -> dict:
    ux, uy = float(freestream[0]), float(freestream[1])
    mag = math.hypot(ux, uy)
    ex, ey = (ux / mag, uy / mag) if mag > 0.0 else (1.0, 0.0)
    u = mag if speed is None else speed
    drag = fx * ex + fy * ey
    lift = -fx * ey + fy * ex
    q = 0.5 * rho * u * u * chord
    return {"force_x": fx, "force_y": fy, "drag": drag, "lift": lift,
            "drag_coefficient": drag / q if q > 0.0 else 0.0, "lift_coefficient": lift / q if q > 0.0 else 0.0}
