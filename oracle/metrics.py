"""Run-level diagnostics restated from fb/core/metrics.py:24-104, fb/core/interpolation.py:9-79,
fb/solvers/lbm.py:1054-1101 (cut-link adjacent normal speed) and fb/solvers/pic_flip.py:1117-1136 (particle
census), as NumPy on host copies of a solver's fields (test infrastructure: the checker of the device
diagnostics, never the product).

Arithmetic follows NumPy's in the field's dtype, as the reference's does: np.gradient quotients and squares
in float32 for a float32 field, means by np.mean.

Pinned by (tests/test_oracle_metrics.py): closed forms (uniform flow energy, solid-body rotation enstrophy
2 Omega^2, the divergence of a linear field, exact bilinear reproduction of linear fields, hand-counted wake
rows and recirculating cells, a hand-computed cut link with a known wall fraction and rotation, the periodic
exclusion of wrapped links) and the census of hand-placed particles.
"""

from __future__ import annotations

import numpy as np


def _dx(dom):
    return (dom.bounds[0][1] - dom.bounds[0][0]) / dom.nx


def _dy(dom):
    return (dom.bounds[1][1] - dom.bounds[1][0]) / dom.ny


def _periodic(dom, axis):
    return axis in dom.periodic_axes


def kinetic_energy(velocity):
    """fb/core/metrics.py:24-31: 0.5 mean of the component-summed square."""
    return 0.5 * float(np.mean(np.sum(velocity * velocity, axis=-1)))


def vorticity(velocity, dom):
    """fb/core/metrics.py:40-43: dv/dx - du/dy by np.gradient (central inside, one-sided at the ends)."""
    return np.gradient(velocity[:, :, 1], _dx(dom), axis=1) - np.gradient(velocity[:, :, 0], _dy(dom), axis=0)


def enstrophy(velocity, dom):
    """fb/core/metrics.py:46-48."""
    w = vorticity(velocity, dom)
    return 0.5 * float(np.mean(w * w))


def divergence_l2(velocity, dom):
    """fb/core/metrics.py:51-55."""
    div = np.gradient(velocity[:, :, 0], _dx(dom), axis=1) + np.gradient(velocity[:, :, 1], _dy(dom), axis=0)
    return float(np.sqrt(np.mean(div * div)))


def _centres_x(dom):
    x0, x1 = dom.bounds[0]
    return np.linspace(x0 + 0.5 * _dx(dom), x1 - 0.5 * _dx(dom), dom.nx)


def wake_width(velocity, dom, pivot_x, chord=1.0, freestream_u=1.0, solid=None, threshold=0.1):
    """fb/core/metrics.py:64-87: rows holding a fluid cell beyond pivot_x + chord whose streamwise deficit
    exceeds the threshold, times dy."""
    cols = _centres_x(dom) > pivot_x + chord
    if not cols.any():
        return 0.0
    deficit = freestream_u - velocity[:, cols, 0] > threshold * abs(freestream_u)
    if solid is not None:
        deficit &= ~solid[:, cols]
    return float(np.count_nonzero(deficit.any(axis=1)) * _dy(dom))


def recirculation_area(velocity, dom, pivot_x, solid=None):
    """fb/core/metrics.py:90-104: fluid cells downstream of the pivot with u < 0, times dx dy."""
    cols = _centres_x(dom) > pivot_x
    back = velocity[:, cols, 0] < 0.0
    if solid is not None:
        back &= ~solid[:, cols]
    return float(np.count_nonzero(back) * _dx(dom) * _dy(dom))


def sample_scalar(field, points, dom, offset=(0.5, 0.5)):
    """fb/core/interpolation.py:9-52: bilinear with the index wrapped on a periodic axis and the coordinate
    clipped to the field otherwise; the result in the field's dtype."""
    gx = (points[:, 0] - dom.bounds[0][0]) / _dx(dom) - offset[0]
    gy = (points[:, 1] - dom.bounds[1][0]) / _dy(dom) - offset[1]
    px, py = _periodic(dom, "x"), _periodic(dom, "y")
    gx = np.mod(gx, dom.nx) if px else np.clip(gx, 0.0, field.shape[1] - 1.0)
    gy = np.mod(gy, dom.ny) if py else np.clip(gy, 0.0, field.shape[0] - 1.0)
    i0 = np.floor(gx).astype(np.int64)
    j0 = np.floor(gy).astype(np.int64)
    tx, ty = gx - i0, gy - j0
    i1 = np.mod(i0 + 1, dom.nx) if px else np.minimum(i0 + 1, field.shape[1] - 1)
    j1 = np.mod(j0 + 1, dom.ny) if py else np.minimum(j0 + 1, field.shape[0] - 1)
    out = ((1.0 - tx) * (1.0 - ty) * field[j0, i0] + tx * (1.0 - ty) * field[j0, i1]
           + (1.0 - tx) * ty * field[j1, i0] + tx * ty * field[j1, i1])
    return np.asarray(out, dtype=field.dtype)


def sample_vector(field, points, dom):
    """fb/core/interpolation.py:67-79."""
    return np.stack([sample_scalar(field[:, :, k], points, dom) for k in range(field.shape[2])], axis=1)


# D2Q9 link directions other than rest (the set of fb/solvers/lbm.py's velocity table)
_LINKS = ((1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1))


def cut_link_adjacent_normal_speed(velocity, solid, sdf, centres, dom, pivot, omega_rad):
    """fb/solvers/lbm.py:1054-1101: over links from a fluid cell into a solid one (wrapped links excluded on a
    non-periodic axis), the wall point at the clipped signed-distance fraction of the link, and the largest
    |(u - omega x (p - pivot)) . link| / |link|."""
    dx, dy = _dx(dom), _dy(dom)
    best = 0.0
    for cx, cy in _LINKS:
        dest = np.roll(solid, shift=(-cy, -cx), axis=(0, 1))
        link = ~solid & dest
        if not _periodic(dom, "x"):
            if cx > 0:
                link[:, -cx:] = False
            elif cx < 0:
                link[:, :-cx] = False
        if not _periodic(dom, "y"):
            if cy > 0:
                link[-cy:, :] = False
            elif cy < 0:
                link[:-cy, :] = False
        if not link.any():
            continue
        dd = np.roll(sdf, shift=(-cy, -cx), axis=(0, 1))
        frac = np.clip(sdf / np.maximum(sdf - dd, 1.0e-12), 0.05, 1.0)[link]
        lx, ly = cx * dx, cy * dy
        pts = centres[link].copy()
        pts[:, 0] += frac * lx
        pts[:, 1] += frac * ly
        rel_p = pts - np.asarray(pivot, dtype=pts.dtype)
        wall = np.stack((-omega_rad * rel_p[:, 1], omega_rad * rel_p[:, 0]), axis=1)
        rel = velocity[link] - wall
        proj = np.abs((rel[:, 0] * lx + rel[:, 1] * ly) / float(np.hypot(lx, ly)))
        best = max(best, float(np.max(proj)))
    return best


def particle_census(positions, solid, dom):
    """fb/solvers/pic_flip.py:1117-1136: particles per fluid cell (fb/core/particle_population.py:19-41)."""
    from .pic import particle_cell_counts
    from .mac import Dom

    d = Dom(dom.bounds, dom.nx, dom.ny, tuple(dom.periodic_axes))
    fc = particle_cell_counts(positions, d).reshape(dom.ny, dom.nx)[~solid]
    return {
        "empty_fluid_cell_fraction": float(np.count_nonzero(fc == 0)) / fc.size,
        "underfilled_fluid_cell_fraction": float(np.count_nonzero(fc < 4)) / fc.size,
        "p05_particles_per_fluid_cell": float(np.percentile(fc, 5.0)),
        "p95_particles_per_fluid_cell": float(np.percentile(fc, 95.0)),
        "max_particles_per_fluid_cell": float(np.max(fc)),
    }
