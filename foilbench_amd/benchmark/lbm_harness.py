"""Kernel-level LBM harness: one lattice with fixed TRT parameters, stepped on the GPU.

This is the configuration the headline cell-steps/s figure is quoted on: a lattice driven at
an explicit lattice speed (above the solver lifecycle's Mach cap of 0.08, so it bypasses the
solver's own time mapping, fb/solvers/lbm.py:370-410) with the solver's own kernels.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .. import _native as N
from ..core.geometry import domain_struct, foil_struct, pose_struct
from ..core.models import DomainSpec, FoilSpec
from ..solvers.lbm import feq9, sponge_profiles

# Algorithmic HBM bytes per cell and lattice step: 9 populations read + 9 written.
BYTES_PER_CELL = {N.FB_F64: 144, N.FB_F32: 72, N.FB_F32_MIXED: 72, N.FB_F16_MIXED: 36}


@dataclass(frozen=True)
class LatticeCase:
    nx: int = 1024
    ny: int = 512
    bounds: tuple = ((-1.6, 4.8), (-1.6, 1.6))
    naca: str = "2412"
    chord: float = 1.0
    pivot: tuple = (0.0, 0.0)
    angle_degrees: float = -12.0
    u_lattice: float = 0.05
    reynolds: float = 1000.0
    magic: float = 3.0 / 16.0
    channel: int = 0  # bounce-back rows at the top and bottom instead of Zou-He rims

    @property
    def name(self) -> str:
        return (f"lbm-d2q9 TRT NACA{self.naca} AoA{-self.angle_degrees:g} {self.nx}x{self.ny} "
                f"chord{int(round(self.chord / ((self.bounds[0][1] - self.bounds[0][0]) / self.nx)))} "
                f"u{self.u_lattice:g} Re{self.reynolds:g}")

    def domain(self) -> DomainSpec:
        return DomainSpec(2, self.bounds, (self.nx, self.ny))

    def relaxation(self) -> tuple[float, float, float]:
        """nu = u * chord_cells / Re; tau+ = 1/2 + 3 nu; tau- = 1/2 + magic/(tau+ - 1/2)."""
        dx = (self.bounds[0][1] - self.bounds[0][0]) / self.nx
        nu = self.u_lattice * (self.chord / dx) / self.reynolds
        tp = 0.5 + 3.0 * nu
        tm = 0.5 + self.magic / (tp - 0.5)
        return nu, 1.0 / tp, 1.0 / tm


class Lattice:
    """A device lattice (fb_lbm handle) initialised at the uniform inflow equilibrium."""

    def __init__(self, case: LatticeCase, precision: int, ctx: N.Context | None = None) -> None:
        self.case = case
        self.precision = precision
        self.ctx = ctx or N.Context.default()
        lib = self.ctx.lib
        dom = case.domain()
        sy, sx = sponge_profiles(dom.nx, dom.ny, False, False, bool(case.channel))
        h = C.c_void_p()
        foil = FoilSpec(case.naca, case.chord, case.pivot)
        N.check(lib.fb_lbm_create(self.ctx.handle, C.byref(domain_struct(dom)), C.byref(foil_struct(foil)),
                                  precision, int(case.channel), N.ptr(sy), N.ptr(sx), C.byref(h)))
        self.handle = h
        N.check(lib.fb_lbm_set_pose(h, C.byref(pose_struct(case.angle_degrees))))
        hd = np.float64 if precision == N.FB_F64 else np.float32
        rho = np.ones((dom.ny, dom.nx), dtype=hd)
        vel = np.zeros((dom.ny, dom.nx, 2), dtype=hd)
        vel[..., 0] = case.u_lattice
        N.check(lib.fb_lbm_set_equilibrium(h, N.ptr(rho), N.ptr(vel)))
        self.params = self._params(hd)

    def _params(self, hd) -> N.fb_lbm_params:
        nu, op, om = self.case.relaxation()
        p = N.fb_lbm_params()
        p.omega_plus = op
        p.omega_minus = om
        p.u_lattice = self.case.u_lattice
        t = np.asarray([1.0, 0.0], dtype=hd) * self.case.u_lattice
        p.target_ux = float(t[0])
        p.target_uy = float(t[1])
        p.wall_omega_rad = 0.0
        p.wall_scale = self.case.u_lattice
        p.velocity_scale = 1.0 / self.case.u_lattice
        # shifted storage: the equilibrium in FP64 (the deviation f - w is what the lattice keeps)
        f = feq9(t[0], t[1], np.float64 if self.precision in (N.FB_F32_MIXED, N.FB_F16_MIXED) else hd)
        for d in range(9):
            p.feq_boundary[d] = float(f[d])
            p.feq_uncovered[d] = float(f[d])
        return p

    @property
    def cells(self) -> int:
        return self.case.nx * self.case.ny

    def run(self, steps: int) -> None:
        N.check(self.ctx.lib.fb_lbm_run_fixed(self.handle, C.byref(self.params), steps))

    def advance(self, steps: int) -> N.fb_lbm_stats:
        """`steps` lattice steps at the case's fixed pose, with the postcondition reductions of the
        last step's pre-collision state."""
        poses = (N.fb_pose * steps)(*[pose_struct(self.case.angle_degrees)] * steps)
        st = N.fb_lbm_stats()
        N.check(self.ctx.lib.fb_lbm_advance(self.handle, C.byref(self.params), steps, poses, C.byref(st)))
        return st

    def state(self) -> tuple[np.ndarray, np.ndarray]:
        """Pre-collision populations (ny, nx, 9) and outlet history (ny, 9) after the last advance."""
        hd = np.float64 if self.precision == N.FB_F64 else np.float32
        f = np.empty((self.case.ny, self.case.nx, 9), dtype=hd)
        out = np.empty((self.case.ny, 9), dtype=hd)
        N.check(self.ctx.lib.fb_lbm_get_state(self.handle, N.ptr(f), N.ptr(out)))
        return f, out

    def time(self, steps: int) -> float:
        """Milliseconds of `steps` back-to-back lattice steps, from one HIP event pair on the stream
        the step kernel runs on."""
        ms = C.c_double()
        N.check(self.ctx.lib.fb_lbm_time_fixed(self.handle, C.byref(self.params), steps, C.byref(ms)))
        return ms.value

    def velocity(self) -> np.ndarray:
        hd = np.float64 if self.precision == N.FB_F64 else np.float32
        rho = np.empty((self.case.ny, self.case.nx), dtype=hd)
        vel = np.empty((self.case.ny, self.case.nx, 2), dtype=hd)
        N.check(self.ctx.lib.fb_lbm_macroscopic(self.handle, 1.0, N.ptr(rho), N.ptr(vel)))
        return vel

    def solid(self) -> np.ndarray:
        s = np.empty((self.case.ny, self.case.nx), dtype=np.uint8)
        N.check(self.ctx.lib.fb_lbm_get_geometry(self.handle, None, N.ptr(s)))
        return s.astype(bool)

    def close(self) -> None:
        if self.handle is not None:
            self.ctx.lib.fb_lbm_destroy(self.handle)
            self.handle = None

    def __del__(self) -> None:  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass
