"""Blended PIC/FLIP on the GPU, behind the reference's solver lifecycle.

Host orchestration mirrors ``PicFlipSolver`` (fb/solvers/pic_flip.py:135-1169): option
validation (:248-288), substep planning (:701-741), the transactional advance with the
deterministic CFL retry (:637-677, :885-949), postconditions (:846-930) and evidence
(:960-1008). Each substep (collisions, sort-by-cell P2G, implicit viscosity on faces,
projection, grid update, G2P blend, midpoint advection with swept collisions, wrap and
respawn) runs on the device in csrc/pic.hip and csrc/mac.hip. Population maintenance
(:620-635, every ``pic_population_interval`` advances) is host logic between advances.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _native as N
from ._diagnostics import DiagParams, fluid_counts_from_histogram, sample_points
from ..core.forces import lift_drag
from ..core.geometry import NacaFoil, cell_centers, domain_struct, foil_struct, pose_struct
from ..core.models import (CanonicalFlowState, ControlState, Diagnostics, ImportOutcome, ImportReport,
                           InteractiveTuning, NumericalFailure, RestartState, ReynoldsOutcome, Scenario, SolverInfo,
                           StepReport)
from ..core.population import maintain_particle_population
from ..core.rng import PCG32
from ..core.validation import validate_advance_request, validate_canonical_import, validate_restart_state
from .stable_fluids import initial_velocity

_SEED_STREAM = 71  # fb/solvers/pic_flip.py:279


def _float_option(sc: Scenario, name: str, default: float) -> float:
    v = sc.solver_options.get(name, default)
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        raise TypeError(f"{name} must be numeric")
    return float(v)


def _int_option(sc: Scenario, name: str, default: int) -> int:
    v = sc.solver_options.get(name, default)
    if isinstance(v, bool) or not isinstance(v, int):
        raise TypeError(f"{name} must be an integer")
    return v


def _postcondition_limit(sc: Scenario, name: str) -> float:
    v = sc.solver_options.get(name)
    if v is None:
        return float("inf")
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        raise TypeError(f"{name} must be numeric")
    f = float(v)
    if not np.isfinite(f) or f < 0.0:
        raise ValueError(f"{name} must be finite and non-negative")
    return f


def sweep_table(start_angle: float, angle: float) -> np.ndarray:
    """(cos, sin) of every swept-collision sample angle (fb/solvers/pic_flip.py:565-569) for
    sample counts 2..16, evaluated with NumPy like NacaFoil._to_local."""
    rows = []
    for count in range(2, 17):
        for s in range(1, count + 1):
            fraction = s / count
            a = np.deg2rad(start_angle + fraction * (angle - start_angle))
            rows.append((float(np.cos(a)), float(np.sin(a))))
    return np.ascontiguousarray(np.asarray(rows, dtype=np.float64))


class PicFlipSolver:
    info = SolverInfo(
        id="pic-flip",
        display_name="Blended PIC/FLIP",
        dimensions=(2,),
        supports_moving_boundary=True,
        supported_precisions=("float32", "float64"),
        acceleration="gfx950 HIP sort-by-cell P2G/G2P + MAC grid (foilbench_amd)",
    )

    def __init__(self, form: str = "plain", ctx: N.Context | None = None, particles_per_cell: int = 4) -> None:
        if form not in ("plain", "mixed"):
            raise ValueError("PIC/FLIP forms: 'plain' (FP64 pressure) or 'mixed' (FP32 CG + FP64 refinement)")
        if not 1 <= int(particles_per_cell) <= 64:
            raise ValueError("particles_per_cell must be in [1, 64]")
        self._form = form
        self._ctx = ctx
        self._per_cell = int(particles_per_cell)
        self._handle = None
        self._mac = None
        self._scenario: Scenario | None = None
        self._geometry: NacaFoil | None = None
        self._control = ControlState(0.0, 0.0, 0.0)
        self._time = 0.0
        self._blend = 0.95
        self._settling_steps = 0
        self._reynolds = 1.0
        self._projection_warning = ""
        self._last_projection = ("relative-l2", 0.0, 0, 0.0, True)
        self._last_viscosity = ("update-linf", 0.0, 0, 0.0, True)
        self._reseeded_last_step = 0
        self._advance_count = 0
        self._population_interval = 8
        self._cfl = 0.75
        self._swept_collisions_last_step = 0
        self._revision = 0
        self._unsupported_face_fraction = 0.0
        self._native_divergence_linf: float | None = None
        self._native_solid_leakage: float | None = None

    # ------------------------------------------------------------------ properties
    @property
    def reynolds(self) -> float:
        return self._reynolds

    @property
    def state_revision(self) -> int:
        return self._revision

    def set_reynolds(self, reynolds: float) -> ReynoldsOutcome:
        if not np.isfinite(reynolds) or reynolds <= 0.0:
            raise ValueError("Reynolds number must be finite and positive")
        r = float(reynolds)
        if r != self._reynolds:
            self._revision += 1
        self._reynolds = r
        return ReynoldsOutcome(r, r)

    @property
    def blend(self) -> float:
        return self._blend

    @blend.setter
    def blend(self, value: float) -> None:
        if not np.isfinite(value):
            raise ValueError("PIC/FLIP blend must be finite")
        b = float(np.clip(value, 0.0, 1.0))
        if b != self._blend and self._scenario is not None:
            self._revision += 1
        self._blend = b

    def interactive_tuning(self) -> InteractiveTuning:
        b = self._blend
        return InteractiveTuning("pic-flip-blend", "blend", b, f"{b:.2f}", b > 0.0, b < 1.0)

    def adjust_interactive_tuning(self, direction: int) -> InteractiveTuning:
        self.blend = self._blend + (-0.05 if direction < 0 else 0.05)
        return self.interactive_tuning()

    def apply_interactive_tuning(self, value) -> InteractiveTuning:
        if not isinstance(value, (int, float)):
            raise TypeError("PIC/FLIP tuning value must be numeric")
        self.blend = float(value)
        return self.interactive_tuning()

    # ------------------------------------------------------------------ plumbing
    def _lib(self):
        return self._ctx.lib

    def _require(self) -> tuple[Scenario, NacaFoil]:
        if self._scenario is None or self._geometry is None or self._handle is None:
            raise RuntimeError("solver has not been initialized")
        return self._scenario, self._geometry

    def _dtype(self):
        return self._scenario.dtype

    def _create_handle(self) -> None:
        if self._handle is not None:
            self._lib().fb_pic_destroy(self._handle)
            self._handle = None
        sc = self._scenario
        prec = N.FB_F32 if sc.precision == "float32" else N.FB_F64
        h = C.c_void_p()
        N.check(self._lib().fb_pic_create(self._ctx.handle, C.byref(domain_struct(sc.domain)),
                                          C.byref(foil_struct(sc.foil)), prec, C.byref(h)))
        self._handle = h
        m = C.c_void_p()
        N.check(self._lib().fb_pic_mac(h, C.byref(m)))
        self._mac = m

    def _set_pose(self, angle: float) -> None:
        N.check(self._lib().fb_pic_set_pose(self._handle, C.byref(pose_struct(angle))))

    def _set_grid(self, vel: np.ndarray) -> None:
        dt = self._dtype()
        N.check(self._lib().fb_pic_set_grid(self._handle, N.ptr(np.ascontiguousarray(vel[..., 0], dtype=dt)),
                                            N.ptr(np.ascontiguousarray(vel[..., 1], dtype=dt))))

    def grid_velocity(self) -> np.ndarray:
        sc, _ = self._require()
        dt = self._dtype()
        gu = np.empty((sc.domain.ny, sc.domain.nx), dt)
        gv = np.empty_like(gu)
        N.check(self._lib().fb_pic_get_grid(self._handle, N.ptr(gu), N.ptr(gv)))
        return np.stack((gu, gv), axis=-1)

    def particles(self) -> tuple[np.ndarray, np.ndarray]:
        self._require()
        n = C.c_int64()
        N.check(self._lib().fb_pic_count(self._handle, C.byref(n)))
        pos = np.empty((n.value, 2), self._dtype())
        vel = np.empty_like(pos)
        N.check(self._lib().fb_pic_get_particles(self._handle, N.ptr(pos), N.ptr(vel)))
        return pos, vel

    def _set_particles(self, pos: np.ndarray, vel: np.ndarray) -> None:
        dt = self._dtype()
        pos = np.ascontiguousarray(pos, dtype=dt)
        vel = np.ascontiguousarray(vel, dtype=dt)
        N.check(self._lib().fb_pic_set_particles(self._handle, pos.shape[0], N.ptr(pos), N.ptr(vel)))

    def solid_mask(self) -> np.ndarray:
        sc, _ = self._require()
        s = np.empty((sc.domain.ny, sc.domain.nx), dtype=np.uint8)
        N.check(self._lib().fb_mac_get_geometry(self._mac, None, N.ptr(s)))
        return s.astype(bool)

    def faces(self) -> tuple[np.ndarray, np.ndarray]:
        """Faces of the last projection."""
        sc, _ = self._require()
        dt = self._dtype()
        u = np.empty((sc.domain.ny, sc.domain.nx + 1), dt)
        v = np.empty((sc.domain.ny + 1, sc.domain.nx), dt)
        N.check(self._lib().fb_mac_get_faces(self._mac, N.ptr(u), N.ptr(v)))
        return u, v

    def rng_words(self) -> tuple[int, int]:
        s = C.c_uint64()
        i = C.c_uint64()
        N.check(self._lib().fb_pic_get_rng(self._handle, C.byref(s), C.byref(i)))
        return int(s.value), int(i.value)

    def _set_rng(self, words: tuple[int, int]) -> None:
        N.check(self._lib().fb_pic_set_rng(self._handle, words[0], words[1]))

    def _stats(self, angle: float) -> N.fb_pic_stats:
        st = N.fb_pic_stats()
        N.check(self._lib().fb_pic_get_stats(self._handle, C.byref(pose_struct(angle)), C.byref(st)))
        return st

    def _collide(self, control: ControlState) -> None:
        N.check(self._lib().fb_pic_resolve_collisions(self._handle, C.byref(pose_struct(control.angle_degrees)),
                                                      float(np.deg2rad(control.angular_velocity_degrees))))

    def _seed(self) -> None:
        n = C.c_int64()
        N.check(self._lib().fb_pic_seed(self._handle, self._per_cell, C.byref(n)))

    def _particle_to_grid(self) -> None:
        frac = C.c_double()
        N.check(self._lib().fb_pic_particle_to_grid(self._handle, 0, C.byref(frac)))
        self._unsupported_face_fraction = float(frac.value)

    # ------------------------------------------------------------------ lifecycle
    def initialize(self, scenario: Scenario, geometry: NacaFoil, seed: int) -> None:
        # fb/solvers/pic_flip.py:248-288
        if scenario.domain.dimension != 2:
            raise NotImplementedError("PIC/FLIP Phase 1 supports only 2D")
        if self._ctx is None:
            self._ctx = N.Context.default()
        self._scenario = scenario
        self._geometry = geometry if isinstance(geometry, NacaFoil) else NacaFoil(geometry.spec)
        self._control = scenario.control_at(0.0)
        self._time = 0.0
        self._reynolds = float(scenario.reynolds)
        self._revision = 0
        self._unsupported_face_fraction = 0.0
        self._native_divergence_linf = None
        self._native_solid_leakage = None
        self._last_projection = ("relative-l2", 0.0, 0, 0.0, True)
        self._last_viscosity = ("update-linf", 0.0, 0, 0.0, True)
        blend = scenario.solver_options.get("pic_flip_blend", 0.95)
        if not isinstance(blend, (int, float)):
            raise TypeError("pic_flip_blend must be numeric")
        self.blend = float(blend)
        interval = scenario.solver_options.get("pic_population_interval", 8)
        if not isinstance(interval, int) or isinstance(interval, bool):
            raise TypeError("pic_population_interval must be an integer")
        if interval < 1:
            raise ValueError("pic_population_interval must be positive")
        self._population_interval = interval
        cfl = scenario.solver_options.get("pic_cfl", 0.75)
        if not isinstance(cfl, (int, float)):
            raise TypeError("pic_cfl must be numeric")
        if not np.isfinite(cfl) or not 0.0 < float(cfl) <= 1.0:
            raise ValueError("pic_cfl must be in (0, 1]")
        self._cfl = float(cfl)
        self._advance_count = 0
        self._settling_steps = 0
        self._create_handle()
        self._set_rng(PCG32(seed, stream=_SEED_STREAM).checkpoint())
        self._set_pose(self._control.angle_degrees)
        self._set_grid(initial_velocity(scenario))
        self._seed()
        # the P2G of initialize only records the unsupported fraction (pic_flip.py:287)
        self._particle_to_grid()
        self._revision = 0

    def restart(self, scenario: Scenario, geometry: NacaFoil, seed: int, start: RestartState) -> None:
        # fb/solvers/pic_flip.py:290-315
        validate_restart_state(start)
        self.initialize(scenario, geometry, seed)
        self.set_reynolds(start.reynolds)
        self._time = start.time
        self._control = ControlState(start.time, start.angle_degrees, 0.0)
        self._set_pose(start.angle_degrees)
        solid = self.solid_mask()
        vel = initial_velocity(scenario)
        wall = self._wall_grid(self._control)
        vel[solid] = wall[solid]
        self._set_grid(vel)
        self._seed()
        self._particle_to_grid()
        self._revision = 0
        self._unsupported_face_fraction = 0.0

    def _wall_grid(self, control: ControlState) -> np.ndarray:
        sc, geometry = self._require()
        pts = cell_centers(sc.domain).reshape(-1, 2)
        return geometry.wall_velocity(pts, control).reshape(sc.domain.ny, sc.domain.nx, 2)

    def _params(self, dt: float, sub: ControlState, start: ControlState) -> N.fb_pic_params:
        sc, geometry = self._require()
        p = N.fb_pic_params()
        p.dt = dt
        p.viscosity = sc.reference_speed * sc.foil.chord / self._reynolds
        p.tolerance = _float_option(sc, "pressure_tolerance", 1.0e-5)
        p.max_iterations = _int_option(sc, "pressure_max_iterations", 640)
        p.channel_walls = 1 if str(sc.solver_options.get("initial_condition", "")) == "poiseuille" else 0
        p.freestream_x = float(sc.freestream[0])
        p.freestream_y = float(sc.freestream[1])
        p.wall_omega_rad = float(np.deg2rad(sub.angular_velocity_degrees))
        p.blend = 0.0 if self._settling_steps > 0 else self._blend
        p.wall_travel = float(abs(np.deg2rad(sub.angle_degrees - start.angle_degrees)) * geometry.maximum_radius)
        p.maximum_radius = float(geometry.maximum_radius)
        p.mixed_poisson = 1 if self._form == "mixed" else 0
        return p

    def _raise_grid_status(self, rep: N.fb_sf_report) -> None:
        st = int(rep.status)
        if st == 0:
            return
        if st == 1:
            raise NumericalFailure("convergence_failure", "PIC/FLIP implicit viscosity did not converge", "viscosity",
                                   {"criterion": "update-linf", "iterations": int(rep.viscosity_iterations),
                                    "tolerance": self._last_viscosity[1],
                                    "final_residual": float(rep.viscosity_residual)})
        if st == 8:
            raise NumericalFailure("projection_failure", "PIC/FLIP pressure solve did not converge", "projection",
                                   {"criterion": "relative-l2", "iterations": int(rep.pressure_iterations),
                                    "tolerance": self._last_projection[1],
                                    "relative_residual": float(rep.pressure_residual)})
        reason, stage, msg = N.SF_STATUS[st]
        raise NumericalFailure(reason, f"PIC/FLIP {msg}", stage)

    def advance(self, control: ControlState, target_dt: float) -> StepReport:
        # fb/solvers/pic_flip.py:637-677
        minimum = 1
        retries = 0
        while True:
            try:
                return self._advance_once(control, target_dt, minimum, retries)
            except NumericalFailure as err:
                acc = err.evidence.get("accepted_cfl")
                mx = err.evidence.get("maximum_cfl")
                att = err.evidence.get("substeps")
                if not (err.reason == "stability_limit" and isinstance(acc, float) and isinstance(mx, float)
                        and isinstance(att, int) and acc > mx > 0.0):
                    raise
                nxt = max(att + 1, int(np.ceil(1.05 * att * acc / mx)))
                if nxt > 512:
                    raise NumericalFailure("stability_limit", "PIC/FLIP retry requires too many internal substeps",
                                           "particle-advection", {**err.evidence, "required_substeps": nxt,
                                                                  "maximum_substeps": 512,
                                                                  "stability_retries": retries}) from err
                retries += 1
                minimum = nxt

    def _advance_once(self, control: ControlState, target_dt: float, minimum_substeps: int,
                      retries: int) -> StepReport:
        # fb/solvers/pic_flip.py:679-1008
        sc, geometry = self._require()
        validate_advance_request(self._time, control, target_dt, sc.precision)
        st0 = self._stats(self._control.angle_degrees)
        if st0.nonfinite:
            raise NumericalFailure("nonfinite_state", "PIC/FLIP input state is non-finite", "postcondition")
        max_speed = max(st0.max_grid_speed, st0.max_particle_speed, abs(sc.freestream[0]), 1.0e-6)
        bav = control.angular_velocity_degrees
        sweep_av = (control.angle_degrees - self._control.angle_degrees) / target_dt
        wall_speed = geometry.maximum_radius * max(abs(np.deg2rad(bav)), abs(np.deg2rad(sweep_av)))
        transport = max(max_speed, wall_speed)
        spacing = min(sc.domain.dx, sc.domain.dy)
        stable_dt = self._cfl * spacing / max(transport, 1.0e-6)
        substeps = max(minimum_substeps, int(np.ceil(target_dt / stable_dt)))
        if substeps > 512:
            raise NumericalFailure("stability_limit", "PIC/FLIP motion requires too many internal substeps",
                                   "particle-advection", {"required_substeps": substeps, "maximum_substeps": 512,
                                                          "maximum_particle_speed": max_speed,
                                                          "maximum_wall_speed": wall_speed})
        dt = target_dt / substeps
        N.check(self._lib().fb_pic_checkpoint(self._handle))
        saved = (self._control, self._time, self._projection_warning, self._last_projection, self._last_viscosity,
                 self._reseeded_last_step, self._swept_collisions_last_step, self._advance_count,
                 self._settling_steps, self._revision, self._unsupported_face_fraction)
        start_time = self._time
        start_angle = self._control.angle_degrees
        lib = self._lib()
        try:
            self._reseeded_last_step = 0
            self._swept_collisions_last_step = 0
            for k in range(substeps):
                frac = (k + 1) / substeps
                sub = ControlState(start_time + frac * target_dt,
                                   start_angle + frac * (control.angle_degrees - start_angle), bav)
                start_control = self._control
                params = self._params(dt, sub, start_control)
                table = sweep_table(start_control.angle_degrees, sub.angle_degrees)
                rep = N.fb_pic_report()
                N.check(lib.fb_pic_substep(self._handle, C.byref(params), C.byref(pose_struct(sub.angle_degrees)),
                                           N.ptr(table), C.byref(rep)))
                g = rep.grid
                tol = params.tolerance
                self._last_viscosity = ("update-linf", tol, int(g.viscosity_iterations),
                                        float(g.viscosity_residual), bool(g.viscosity_converged))
                if g.status not in (1, 2, 3):
                    self._last_projection = ("relative-l2", tol, int(g.pressure_iterations),
                                             float(g.pressure_residual), bool(g.pressure_converged))
                self._raise_grid_status(g)
                total_faces = (sc.domain.ny * (sc.domain.nx + 1)) + ((sc.domain.ny + 1) * sc.domain.nx)
                self._unsupported_face_fraction = int(rep.unsupported_faces) / total_faces
                self._swept_collisions_last_step += int(rep.swept_hits)
                self._projection_warning = ""
                self._control = sub
            if self._settling_steps > 0:
                self._settling_steps -= 1
            self._advance_count += 1
            if self._advance_count % self._population_interval == 0:
                self._maintain_population(control)
            self._collide(control)
            st = self._stats(control.angle_degrees)
            if st.nonfinite:
                raise NumericalFailure("nonfinite_state", "PIC/FLIP produced non-finite state", "postcondition")
            if st.particles_inside:
                raise NumericalFailure("postcondition_failure", "PIC/FLIP left particles inside the moving solid",
                                       "postcondition", {"unresolved_solid_particles": int(st.particles_inside)})
            report_speed = max(transport, st.max_grid_speed, st.max_particle_speed)
            if not np.isfinite(report_speed):
                raise NumericalFailure("nonfinite_state", "PIC/FLIP produced a non-finite step report",
                                       "postcondition")
            accepted = dt * report_speed / spacing
            if accepted > self._cfl * (1.0 + 1.0e-6):
                raise NumericalFailure("stability_limit", "PIC/FLIP post-step motion exceeded its swept envelope",
                                       "particle-advection", {"accepted_cfl": accepted, "maximum_cfl": self._cfl,
                                                              "substeps": substeps})
            ms = N.fb_mac_stats()
            N.check(lib.fb_mac_get_stats(self._mac, float(np.deg2rad(control.angular_velocity_degrees)),
                                         C.byref(ms)))
            native_div = float(ms.divergence_linf)
            native_leak = float(ms.solid_leakage)
            div_limit = _postcondition_limit(sc, "mac_maximum_divergence_linf")
            leak_limit = _postcondition_limit(sc, "mac_maximum_solid_leakage")
            if native_div > div_limit or native_leak > leak_limit:
                raise NumericalFailure("postcondition_failure",
                                       "PIC/FLIP exceeded a configured MAC postcondition limit", "postcondition",
                                       {"divergence_linf": native_div, "maximum_divergence_linf": div_limit,
                                        "solid_leakage": native_leak, "maximum_solid_leakage": leak_limit})
        except Exception:
            N.check(lib.fb_pic_rollback(self._handle))
            (self._control, self._time, self._projection_warning, self._last_projection, self._last_viscosity,
             self._reseeded_last_step, self._swept_collisions_last_step, self._advance_count,
             self._settling_steps, self._revision, self._unsupported_face_fraction) = saved
            raise
        self._time = start_time + target_dt
        self._control = ControlState(self._time, control.angle_degrees, control.angular_velocity_degrees)
        self._revision += 1
        self._native_divergence_linf = native_div
        self._native_solid_leakage = native_leak
        fluid = max(1, int(st.fluid_cells))
        warnings = () if not self._projection_warning else (self._projection_warning,)
        lp = self._last_projection
        lv = self._last_viscosity
        cfl_measure = dt * report_speed / spacing
        return StepReport(target_dt, target_dt, substeps, report_speed, warnings, self._revision, {
            "maximum_particle_speed": float(st.max_particle_speed),
            "stability_retries": retries,
            "maximum_wall_speed": wall_speed,
            "maximum_particle_cfl": cfl_measure,
            "maximum_characteristic_displacement": cfl_measure,
            "particle_count": int(st.particle_count),
            "empty_cell_fraction": float(st.empty_cells) / fluid,
            "underfilled_cell_fraction": float(st.underfilled_cells) / fluid,
            "unresolved_solid_particles": 0,
            "unsupported_face_fraction": self._unsupported_face_fraction,
            "pressure_converged": True,
            "pressure_iterations": lp[2],
            "pressure_relative_residual": lp[3],
            "viscosity_converged": lv[4],
            "viscosity_iterations": lv[2],
            "viscosity_final_residual": lv[3],
            "divergence_linf": native_div,
            "solid_leakage": native_leak,
            "requested_reynolds": self._reynolds,
            "effective_reynolds": self._reynolds,
            "degraded_motion": wall_speed == 0.0 and abs(control.angle_degrees - start_angle) > 1.0e-9,
        })

    def _maintain_population(self, control: ControlState) -> None:
        # fb/solvers/pic_flip.py:620-635 on host copies of the device positions and stream
        sc, _ = self._require()
        pos, vel = self.particles()
        rng = PCG32(0)
        rng.restore(self.rng_words())
        moved, _ = maintain_particle_population(pos, self.solid_mask(), sc.domain, rng)
        if moved.size:
            self._set_particles(pos, vel)
            idx = np.ascontiguousarray(moved, dtype=np.int64)
            N.check(self._lib().fb_pic_regrid(self._handle, N.ptr(idx), idx.size))
            self._reseeded_last_step += int(moved.size)
            self._collide(control)
        self._set_rng(rng.checkpoint())

    # ------------------------------------------------------------------ state exchange
    def sample_velocity(self, points: np.ndarray) -> np.ndarray:
        # fb/solvers/pic_flip.py:1010-1012: the grid velocity sampled bilinearly, on the device
        self._require()
        pts = sample_points(points)
        out = np.empty(pts.shape, dtype=self._dtype())
        N.check(self._lib().fb_pic_sample_velocity(self._handle, pts.shape[0], N.ptr(pts), N.ptr(out)))
        return out

    def export_state(self) -> CanonicalFlowState:
        # fb/solvers/pic_flip.py:1014-1033
        sc, _ = self._require()
        vel = self.grid_velocity()
        vel[self.solid_mask()] = 0.0
        return CanonicalFlowState(
            schema_version=2, dimension=2, bounds=sc.domain.bounds, resolution=sc.domain.resolution,
            periodic_axes=sc.domain.periodic_axes, time=self._time, precision=sc.precision,
            angle_degrees=self._control.angle_degrees, angular_velocity_degrees=self._control.angular_velocity_degrees,
            source_language="python", source_solver=self.info.id, velocity=vel[None, ...], geometry=sc.foil,
            producer_execution_target="native",
        )

    def import_state(self, state: CanonicalFlowState, control: ControlState) -> ImportOutcome:
        # fb/solvers/pic_flip.py:1035-1113
        sc, geometry = self._require()
        N.check(self._lib().fb_pic_checkpoint(self._handle))
        saved = (self._control, self._time, self._settling_steps, self._revision, self._unsupported_face_fraction)

        def restore():
            N.check(self._lib().fb_pic_rollback(self._handle))
            (self._control, self._time, self._settling_steps, self._revision,
             self._unsupported_face_fraction) = saved

        try:
            solid = geometry.mask(sc.domain, control.angle_degrees, self._ctx)
            validate_canonical_import(state, sc, control, solid)
            vel = np.asarray(state.velocity[0], dtype=sc.dtype).copy()
            self._time = state.time
            self._control = control
            self._set_pose(control.angle_degrees)
            wall = self._wall_grid(control)
            vel[solid] = wall[solid]
            self._set_grid(vel)
            self._seed()
            self._settling_steps = 1
            self._native_divergence_linf = None
            self._native_solid_leakage = None
        except NumericalFailure as failure:
            restore()
            return ImportOutcome("rejected", failure.reason, warnings=(str(failure),), stage=failure.stage,
                                 evidence=failure.evidence)
        except Exception:
            restore()
            raise
        self._revision += 1
        report = ImportReport(state.source_solver, self.info.id, ("solver particles", "FLIP velocity delta history"),
                              ("The first imported step is PIC-dominant while FLIP history is rebuilt.",))
        return ImportOutcome("accepted", "none", report, report.warnings)

    def forces(self) -> dict:
        """Force on the foil from the MAC grid's pressure and shear (unit density), with drag, lift and
        their coefficients."""
        sc, _ = self._require()
        f = (C.c_double * 2)()
        nu = sc.reference_speed * sc.foil.chord / self._reynolds
        N.check(self._lib().fb_mac_forces(self._mac, nu, float(np.deg2rad(self._control.angular_velocity_degrees)), f))
        return lift_drag(f[0], f[1], sc.freestream[:2], sc.foil.chord)

    def diagnostics(self) -> Diagnostics:
        # fb/solvers/pic_flip.py:1115-1169
        sc, geometry = self._require()
        d = N.fb_diag()
        hist = np.zeros(N.FB_DIAG_HISTOGRAM, np.int64)
        N.check(self._lib().fb_pic_diagnostics(self._handle, DiagParams(sc).ref(), C.byref(d), N.ptr(hist), None))

        def counts():
            c = np.empty((sc.domain.ny, sc.domain.nx), np.int32)
            N.check(self._lib().fb_pic_diagnostics(self._handle, DiagParams(sc).ref(), C.byref(N.fb_diag()),
                                                   N.ptr(np.zeros_like(hist)), N.ptr(c)))
            return c

        # the fluid cells' particle counts (fb/core/particle_population.py:34-41 at the mask), for NumPy's
        # percentiles of the same values
        fc = fluid_counts_from_histogram(hist, counts, self.solid_mask)
        st = self._stats(self._control.angle_degrees)
        values = {
            "time": self._time,
            "requested_reynolds": self._reynolds,
            "kinetic_energy": d.kinetic_energy,
            "enstrophy": d.enstrophy,
            "divergence_l2": d.divergence_l2,
            "particle_count": float(st.particle_count),
            "unsupported_face_fraction": self._unsupported_face_fraction,
            "empty_fluid_cell_fraction": float(d.empty_cells) / d.fluid_cells,
            "underfilled_fluid_cell_fraction": float(d.underfilled_cells) / d.fluid_cells,
            "p05_particles_per_fluid_cell": float(np.percentile(fc, 5.0)),
            "p95_particles_per_fluid_cell": float(np.percentile(fc, 95.0)),
            "max_particles_per_fluid_cell": float(d.max_particles),
            "reseeded_last_step": float(self._reseeded_last_step),
            "swept_collisions_last_step": float(self._swept_collisions_last_step),
            "particles_inside_solid": float(st.particles_inside),
            "wake_width": d.wake_width,
            "recirculation_area": d.recirculation_area,
        }
        if self._native_divergence_linf is not None:
            values["divergence_linf"] = self._native_divergence_linf
        if self._native_solid_leakage is not None:
            values["solid_leakage"] = self._native_solid_leakage
        if not all(np.isfinite(v) for v in values.values()):
            raise NumericalFailure("nonfinite_state", "PIC/FLIP produced non-finite diagnostics")
        warnings = () if not self._projection_warning else (self._projection_warning,)
        return Diagnostics(values, warnings, self._revision)

    def __del__(self) -> None:  # pragma: no cover
        try:
            if self._handle is not None and self._ctx is not None:
                self._ctx.lib.fb_pic_destroy(self._handle)
                self._handle = None
        except Exception:
            pass
