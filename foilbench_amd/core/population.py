"""Deterministic particle population maintenance (fb/core/particle_population.py:19-135), run on
the host between PIC/FLIP advances (every `pic_population_interval` advances). The inputs are
the device positions copied to the host; moved particles are written back by the caller.
"""

from __future__ import annotations

import numpy as np

from .rng import PCG32


def particle_cell_ids(positions: np.ndarray, domain) -> np.ndarray:
    """Clipped cell index of each particle in the positions' dtype (particle_population.py:19-31)."""
    (x0, _), (y0, _) = domain.bounds
    cx = np.clip(np.floor((positions[:, 0] - x0) / domain.dx).astype(np.int64), 0, domain.nx - 1)
    cy = np.clip(np.floor((positions[:, 1] - y0) / domain.dy).astype(np.int64), 0, domain.ny - 1)
    return cy * domain.nx + cx


def particle_cell_counts(positions: np.ndarray, domain) -> np.ndarray:
    return np.bincount(particle_cell_ids(positions, domain), minlength=domain.nx * domain.ny).astype(np.int64)


def maintain_particle_population(positions: np.ndarray, solid: np.ndarray, domain, rng: PCG32, minimum: int = 2,
                                 target: int = 4, maximum: int = 8) -> tuple[np.ndarray, np.ndarray]:
    """Redistribute particles so every fluid cell holds minimum..maximum of them; returns the
    moved particle indices and the final per-cell counts. `positions` is updated in place."""
    if not 0 < minimum <= target <= maximum:
        raise ValueError("particle population limits must satisfy 0 < minimum <= target <= maximum")
    n = positions.shape[0]
    if n == 0:
        raise ValueError("particle population cannot be empty")
    ncell = domain.nx * domain.ny
    ids = particle_cell_ids(positions, domain)
    counts = np.bincount(ids, minlength=ncell).astype(np.int64)
    solid_flat = solid.reshape(-1)
    fluid = np.flatnonzero(~solid_flat)
    if fluid.size == 0:
        raise ValueError("particle population requires at least one fluid cell")
    fc = counts[fluid]
    if np.all((fc >= minimum) & (fc <= maximum)) and not np.any(counts[solid_flat]):
        return np.empty(0, dtype=np.int64), counts

    # target occupancy: solid cells empty, out-of-range fluid cells reset to `target`
    want = np.where(solid_flat, 0, counts)
    want[fluid[(fc < minimum) | (fc > maximum)]] = target
    surplus = n - int(want.sum())
    if surplus != 0:
        # fill (or drain) fluid cells one slot at a time, least (most) occupied first, ties by index
        level = want[fluid]
        if surplus > 0:
            room = maximum - level
            order = np.lexsort((fluid, level))
            step = 1
        else:
            room = level - minimum
            order = np.lexsort((fluid, -level))
            step = -1
        slots = np.repeat(fluid[order], room[order])
        need = abs(surplus)
        if slots.size < need:
            raise RuntimeError("particle count exceeds configured maximum population" if surplus > 0
                               else "particle count cannot satisfy configured minimum population")
        np.add.at(want, slots[:need], step)

    # donors: particles beyond a cell's wanted count, in (cell, particle index) order
    order = np.argsort(ids, kind="stable")
    sorted_ids = ids[order]
    first = np.searchsorted(sorted_ids, sorted_ids, side="left")
    rank = np.arange(n, dtype=np.int64) - first
    donors = order[rank >= want[sorted_ids]]
    deficit = want - np.minimum(counts, want)
    dest = np.repeat(np.arange(ncell, dtype=np.int64), deficit)
    if donors.size != dest.size:
        raise RuntimeError("particle population redistribution is unbalanced")
    if donors.size:
        jit = rng.random((donors.size, 2)).astype(positions.dtype, copy=False)
        (x0, _), (y0, _) = domain.bounds
        positions[donors, 0] = x0 + (dest % domain.nx + 0.4 + 0.2 * jit[:, 0]) * domain.dx
        positions[donors, 1] = y0 + (dest // domain.nx + 0.4 + 0.2 * jit[:, 1]) * domain.dy
    return donors, particle_cell_counts(positions, domain)
