"""Synthetic serving workloads: scenarios of uneven grid size, foil and Reynolds number arriving
as a Poisson process.

The mix draws, per request: a grid from {128x64, 256x128, 512x256, 1024x512}; a NACA four-digit
foil with camber 0-6 %, camber position 20-50 % and thickness 8-18 %; an angle of attack uniform
in [-5, 20] degrees; a Reynolds number log-uniform in [1e2, 1e4]; the Stable Fluids or the D2Q9
solver with equal probability; and a step count uniform in [200, 5000] (scaled by
`step_scale`). Inter-arrival times are exponential with mean 1 / rate.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ..core.models import ControlKeyframe, DomainSpec, FoilSpec, Scenario

GRIDS = ((128, 64), (256, 128), (512, 256), (1024, 512))
SF = "stable-fluids"
LBM = "lbm-d2q9"

# Lattice speed of the serving lattices (lattice units per step): below the solver lifecycle's
# Mach cap, so one output step is one lattice step.
LATTICE_SPEED = 0.04


@dataclass(frozen=True)
class Request:
    id: int
    arrival_s: float
    solver: str
    nx: int
    ny: int
    naca: str
    aoa_degrees: float
    reynolds: float
    steps: int

    @property
    def cells(self) -> int:
        return self.nx * self.ny

    @property
    def cell_steps(self) -> int:
        return self.nx * self.ny * self.steps

    def scenario(self, precision: str) -> Scenario:
        """The request as a scenario of the solver lifecycle. The foil is pitched nose-up by the
        angle of attack (control angle = -AoA). Stable Fluids runs on [-1.5, 2.5] x [-1, 1] with
        output_dt = 0.5 dx / U (CFL 0.5, one substep per step); the lattice runs on
        [-1.6, 4.8] x [-1.6, 1.6] with output_dt = LATTICE_SPEED dx / U (one lattice step)."""
        if self.solver == SF:
            bounds = ((-1.5, 2.5), (-1.0, 1.0))
            dx = 4.0 / self.nx
            out_dt = 0.5 * dx
            opts = {"stable_advection": "maccormack", "pressure_tolerance": 1.0e-6,
                    "pressure_max_iterations": 640, "stable_cfl": 1.0}
        else:
            bounds = ((-1.6, 4.8), (-1.6, 1.6))
            dx = 6.4 / self.nx
            out_dt = LATTICE_SPEED * dx
            opts = {}
        return Scenario(
            schema_version=1, id=f"req-{self.id}", domain=DomainSpec(2, bounds, (self.nx, self.ny)),
            reynolds=self.reynolds, freestream=(1.0, 0.0), foil=FoilSpec(self.naca, 1.0, (0.0, 0.0)),
            controls=(ControlKeyframe(0.0, -self.aoa_degrees),), duration=out_dt * self.steps, output_dt=out_dt,
            precision=precision, seed=0, solver_options=opts)


def naca_code(camber_pct: int, position_tenths: int, thickness_pct: int) -> str:
    if camber_pct == 0:
        return f"00{thickness_pct:02d}"
    return f"{camber_pct}{position_tenths}{thickness_pct:02d}"


def serving_mix(n: int = 256, rate: float = 50.0, seed: int = 2024, step_scale: float = 1.0,
                grids=GRIDS, solvers=(SF, LBM), min_steps: int = 200, max_steps: int = 5000) -> list[Request]:
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(n):
        t += float(rng.exponential(1.0 / rate))
        nx, ny = grids[int(rng.integers(0, len(grids)))]
        code = naca_code(int(rng.integers(0, 7)), int(rng.integers(2, 6)), int(rng.integers(8, 19)))
        aoa = float(rng.uniform(-5.0, 20.0))
        re = float(10.0 ** rng.uniform(2.0, 4.0))
        solver = solvers[int(rng.integers(0, len(solvers)))]
        steps = int(rng.integers(min_steps, max_steps + 1))
        steps = max(2, int(round(steps * step_scale)))
        out.append(Request(k, t, solver, nx, ny, code, aoa, re, steps))
    return out
