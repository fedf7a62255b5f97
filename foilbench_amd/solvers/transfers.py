"""The particle/grid transfer adapter of fb/solvers/_numba_adapter.py:35-160 on the GPU.

Same names, arguments, return values and argument checks as the reference's adapter, each call running the HIP
kernels behind the C ABI (fb_pic_particle_to_cells, fb_pic_cells_to_particle, fb_pic_particle_to_faces,
fb_pic_faces_to_particle) on host arrays:

* particle_to_grid / grid_to_particle: the cell-centred quadratic B-spline pair (kern/pic.py:22-102), the
  stencil wrapped on a periodic axis and clamped to the edge cells otherwise. Both are bit-identical to the
  reference's kernels.
* particle_to_faces / faces_to_particle: the MAC-face pair (kern/pic.py:105-205) that the PIC/FLIP solver runs.
  The gather is bit-identical; the scatter sums each face's contributions in cell order, so FP64 faces agree
  with the reference's particle-order sums to about 1e-13 relative and FP32 faces to about 1e-6.

Two checks are stricter than the reference's: every array of one call must share one dtype (TypeError), and
particle_to_faces needs at least four cells per axis (ValueError). There is no CPU fallback: without the HIP
library or a visible GPU the calls raise.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _native as N


def _require_supported_float(array: np.ndarray, name: str) -> None:
    if array.dtype not in (np.dtype(np.float32), np.dtype(np.float64)):
        raise TypeError(f"{name} must use float32 or float64")


def _precision(dtype: np.dtype) -> int:
    return N.FB_F64 if dtype == np.float64 else N.FB_F32


def _same_dtype(*arrays: np.ndarray) -> None:
    if len({a.dtype for a in arrays}) != 1:
        raise TypeError("transfer arrays must use the same dtype")


def _call(fn, *args) -> None:
    try:
        N.check(fn(*args))
    except N.NativeError as err:
        if err.status == 1:  # FB_ERR_ARG
            raise ValueError(str(err)) from None
        raise


def _c(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a)


def particle_to_grid(positions: np.ndarray, velocities: np.ndarray, x0: float, y0: float, dx: float, dy: float,
                     nx: int, ny: int, freestream: tuple[float, ...], periodic_x: bool = False,
                     periodic_y: bool = False) -> np.ndarray:
    """fb/solvers/_numba_adapter.py:35-71: particle velocities onto a (ny, nx, 2) cell grid."""
    if positions.ndim != 2 or positions.shape[1] != 2:
        raise ValueError("positions must have shape (particle, 2)")
    if velocities.shape != positions.shape:
        raise ValueError("velocities must match positions")
    _require_supported_float(positions, "positions")
    _require_supported_float(velocities, "velocities")
    if positions.dtype != velocities.dtype:
        raise TypeError("positions and velocities must use the same dtype")
    if dx <= 0.0 or dy <= 0.0 or nx <= 0 or ny <= 0:
        raise ValueError("grid spacing and resolution must be positive")
    ctx = N.Context.default()
    pos, vel = _c(positions), _c(velocities)
    grid = np.empty((int(ny), int(nx), 2), dtype=positions.dtype)
    _call(ctx.lib.fb_pic_particle_to_cells, ctx.handle, _precision(pos.dtype), pos.shape[0], N.ptr(pos),
          N.ptr(vel), int(nx), int(ny), float(x0), float(y0), float(dx), float(dy), float(freestream[0]),
          float(freestream[1]), int(bool(periodic_x)), int(bool(periodic_y)), N.ptr(grid))
    return grid


def grid_to_particle(grid: np.ndarray, positions: np.ndarray, x0: float, y0: float, dx: float, dy: float,
                     periodic_x: bool = False, periodic_y: bool = False) -> np.ndarray:
    """fb/solvers/_numba_adapter.py:74-105: the (ny, nx, 2) cell grid sampled at the particles."""
    if grid.ndim != 3 or grid.shape[2] != 2:
        raise ValueError("grid must have shape (ny, nx, 2)")
    if grid.shape[0] == 0 or grid.shape[1] == 0:
        raise ValueError("grid axes must be nonempty")
    if positions.ndim != 2 or positions.shape[1] != 2:
        raise ValueError("positions must have shape (particle, 2)")
    _require_supported_float(grid, "grid")
    _require_supported_float(positions, "positions")
    if grid.dtype != positions.dtype:
        raise TypeError("grid and positions must use the same dtype")
    if dx <= 0.0 or dy <= 0.0:
        raise ValueError("grid spacing must be positive")
    ctx = N.Context.default()
    g, pos = _c(grid), _c(positions)
    out = np.empty_like(pos)
    _call(ctx.lib.fb_pic_cells_to_particle, ctx.handle, _precision(pos.dtype), pos.shape[0], N.ptr(g), N.ptr(pos),
          g.shape[1], g.shape[0], float(x0), float(y0), float(dx), float(dy), int(bool(periodic_x)),
          int(bool(periodic_y)), N.ptr(out))
    return out


def particle_to_faces(positions: np.ndarray, velocities: np.ndarray, fallback_u: np.ndarray,
                      fallback_v: np.ndarray, x0: float, y0: float, dx: float, dy: float, periodic_x: bool = False,
                      periodic_y: bool = False) -> tuple[np.ndarray, np.ndarray, float]:
    """fb/solvers/_numba_adapter.py:108-135: particle velocities onto MAC faces u (ny, nx+1), v (ny+1, nx), with
    the fraction of faces left at their fallback."""
    if positions.ndim != 2 or positions.shape[1] != 2 or velocities.shape != positions.shape:
        raise ValueError("particle positions and velocities must have shape (particle, 2)")
    if fallback_u.shape != (fallback_v.shape[0] - 1, fallback_v.shape[1] + 1):
        raise ValueError("fallback arrays do not form a 2D MAC grid")
    for a, name in ((positions, "positions"), (velocities, "velocities"), (fallback_u, "fallback_u"),
                    (fallback_v, "fallback_v")):
        _require_supported_float(a, name)
    _same_dtype(positions, velocities, fallback_u, fallback_v)
    ctx = N.Context.default()
    pos, vel, fu, fv = _c(positions), _c(velocities), _c(fallback_u), _c(fallback_v)
    ny, nx = fu.shape[0], fv.shape[1]
    u, v = np.empty_like(fu), np.empty_like(fv)
    frac = C.c_double()
    _call(ctx.lib.fb_pic_particle_to_faces, ctx.handle, _precision(pos.dtype), pos.shape[0], N.ptr(pos), N.ptr(vel),
          N.ptr(fu), N.ptr(fv), nx, ny, float(x0), float(y0), float(dx), float(dy), int(bool(periodic_x)),
          int(bool(periodic_y)), N.ptr(u), N.ptr(v), C.byref(frac))
    return u, v, float(frac.value)


def faces_to_particle(u: np.ndarray, v: np.ndarray, positions: np.ndarray, x0: float, y0: float, dx: float,
                      dy: float, periodic_x: bool = False, periodic_y: bool = False) -> np.ndarray:
    """fb/solvers/_numba_adapter.py:138-160: MAC faces sampled at the particles."""
    if u.shape != (v.shape[0] - 1, v.shape[1] + 1):
        raise ValueError("face arrays do not form a 2D MAC grid")
    if positions.ndim != 2 or positions.shape[1] != 2:
        raise ValueError("positions must have shape (particle, 2)")
    for a, name in ((u, "u"), (v, "v"), (positions, "positions")):
        _require_supported_float(a, name)
    _same_dtype(u, v, positions)
    ctx = N.Context.default()
    fu, fv, pos = _c(u), _c(v), _c(positions)
    out = np.empty_like(pos)
    _call(ctx.lib.fb_pic_faces_to_particle, ctx.handle, _precision(pos.dtype), pos.shape[0], N.ptr(fu), N.ptr(fv),
          N.ptr(pos), fv.shape[1], fu.shape[0], float(x0), float(y0), float(dx), float(dy), int(bool(periodic_x)),
          int(bool(periodic_y)), N.ptr(out))
    return out
