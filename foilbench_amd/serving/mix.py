"""Synthetic serving workloads: scenarios of uneven grid size, foil and Reynolds number arriving
as a Poisson process.

The mix draws, per request: a grid from {128x64, 256x128, 512x256, 1024x512}; a NACA four-digit
foil with camber 0-6 %, camber position 20-50 % and thickness 8-18 %; an angle of attack uniform
in [-5, 20] degrees; a Reynolds number log-uniform in [1e2, 1e4]; the Stable Fluids or the D2Q9
solver with equal probability; and a step count uniform in [200, 5000] (scaled by
`step_scale`). Inter-arrival times are exponential with mean 1 / rate. All draws come from the
project's PCG32 stream (seed 42, stream 1001 by default), nine 32-bit words per request, each
read as u = word / 2^32.
"""

from __future__ import annotations

from dataclasses import dataclass

import math

from ..core.models import ControlKeyframe, DomainSpec, FoilSpec, Scenario
from ..core.rng import PCG32

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


def serving_mix(n: int = 256, rate: float = 50.0, seed: int = 42, stream: int = 1001, step_scale: float = 1.0,
                grids=GRIDS, solvers=(SF, LBM), min_steps: int = 200, max_steps: int = 5000) -> list[Request]:
    words = PCG32(seed, stream).uint32(9 * n)
    u = [float(w) / 4294967296.0 for w in words]
    pick = lambda x, k: min(int(x * k), k - 1)  # noqa: E731
    t = 0.0
    out = []
    for k in range(n):
        d = u[9 * k:9 * k + 9]
        t += -math.log1p(-d[0]) / rate
        nx, ny = grids[pick(d[1], len(grids))]
        code = naca_code(pick(d[2], 7), 2 + pick(d[3], 4), 8 + pick(d[4], 11))
        aoa = -5.0 + 25.0 * d[5]
        re = 10.0 ** (2.0 + 2.0 * d[6])
        solver = solvers[pick(d[7], len(solvers))]
        steps = min_steps + pick(d[8], max_steps - min_steps + 1)
        steps = max(2, int(round(steps * step_scale)))
        out.append(Request(k, t, solver, nx, ny, code, aoa, re, steps))
    return out
