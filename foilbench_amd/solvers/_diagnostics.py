"""Host side of the device diagnostics: the parameters of fb/core/metrics.py:64-104 and the lattice cut links
for a scenario, and the census percentiles of fb/solvers/pic_flip.py:1130-1136 from the device histogram."""

from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _native as N


class DiagParams:
    """fb_diag_params of a scenario, with the cell-centre arrays it points to kept alive."""

    def __init__(self, sc, angular_velocity_degrees: float = 0.0) -> None:
        dom = sc.domain
        (x0, x1), (y0, y1) = dom.bounds
        # fb/core/metrics.py:72-77 and fb/core/geometry.py:162-168: cell centres by np.linspace
        self.centers_x = np.ascontiguousarray(np.linspace(x0 + 0.5 * dom.dx, x1 - 0.5 * dom.dx, dom.nx))
        self.centers_y = np.ascontiguousarray(np.linspace(y0 + 0.5 * dom.dy, y1 - 0.5 * dom.dy, dom.ny))
        pivot_x, chord, fs_u = sc.foil.pivot[0], sc.foil.chord, sc.freestream[0]
        p = N.fb_diag_params()
        p.freestream_u = float(fs_u)
        p.wake_threshold = 0.1 * abs(fs_u)
        # centres increase along x, so the columns beyond a coordinate are a suffix
        p.wake_col0 = dom.nx - int(np.count_nonzero(self.centers_x > pivot_x + chord))
        p.down_col0 = dom.nx - int(np.count_nonzero(self.centers_x > pivot_x))
        p.angular_velocity_rad = float(np.deg2rad(angular_velocity_degrees))
        p.diagonal_link = float(np.hypot(dom.dx, dom.dy))
        p.centers_x = self.centers_x.ctypes.data_as(C.POINTER(C.c_double))
        p.centers_y = self.centers_y.ctypes.data_as(C.POINTER(C.c_double))
        self.struct = p

    def ref(self):
        return C.byref(self.struct)


def fluid_counts_from_histogram(hist: np.ndarray, counts_fn, solid_fn) -> np.ndarray:
    """The fluid cells' particle counts in sorted order (what np.percentile and np.max see): expanded from the
    histogram, or read per cell when some cell reaches the histogram's last entry."""
    if hist[-1] == 0:
        return np.repeat(np.arange(hist.size, dtype=np.int64), hist)
    counts = counts_fn().astype(np.int64)
    return np.sort(counts[~solid_fn()])


def sample_points(points: np.ndarray) -> np.ndarray:
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64))
    if pts.ndim != 2 or pts.shape[1] != 2:
        raise ValueError("sample points must match the domain dimension")
    return pts
