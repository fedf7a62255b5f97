"""Stable Fluids on a staggered MAC grid, on the GPU, behind the reference's lifecycle.

Host orchestration mirrors ``StableFluidsSolver`` (fb/solvers/stable_fluids.py:119-738):
substep planning from the maximum cell speed, wall speed and boundary sweep (:335-379),
transactional advance with the deterministic CFL retry (:387-535), MAC postconditions
(:541-567) and evidence (:576-603). Each substep (advection, implicit viscosity, mask,
projection with the masked DST-preconditioned CG) runs in csrc/mac.hip + csrc/poisson.hip.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _native as N
from ._diagnostics import DiagParams, sample_points
from ..core.forces import lift_drag
from ..core.geometry import NacaFoil, cell_centers, domain_struct, foil_struct, pose_struct
from ..core.models import (CanonicalFlowState, ControlState, Diagnostics, ImportOutcome, ImportReport,
                           InteractiveTuning, NumericalFailure, RestartState, ReynoldsOutcome, Scenario, SolverInfo,
                           StepReport)
from ..core.validation import validate_advance_request, validate_canonical_import, validate_restart_state

_MIN_PROJECTION_CFL_LIMIT = 1.0
_MODES = ("maccormack", "semi-lagrangian", "skew-rk2")


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


def initial_velocity(sc: Scenario) -> np.ndarray:
    """fb/solvers/stable_fluids.py:102-116 (scenario set-up on the host)."""
    dom = sc.domain
    vel = np.empty((dom.ny, dom.nx, 2), dtype=sc.dtype)
    vel[...] = np.asarray(sc.freestream[:2], dtype=sc.dtype)
    init = str(sc.solver_options.get("initial_condition", "freestream"))
    pos = cell_centers(dom)
    if init == "taylor-green":
        vel[:, :, 0] = np.sin(pos[:, :, 0]) * np.cos(pos[:, :, 1])
        vel[:, :, 1] = -np.cos(pos[:, :, 0]) * np.sin(pos[:, :, 1])
    elif init == "poiseuille":
        y0, y1 = dom.bounds[1]
        r = 0.5 * (y1 - y0)
        c = 0.5 * (y0 + y1)
        vel[:, :, 0] = 1.5 * (1.0 - ((pos[:, :, 1] - c) / r) ** 2)
        vel[:, :, 1] = 0.0
    return vel


def cell_to_faces(vel: np.ndarray):
    """Initial face values from a host cell field (fb/core/grid.py:32-44)."""
    ny, nx, _ = vel.shape
    u = np.empty((ny, nx + 1), dtype=vel.dtype)
    v = np.empty((ny + 1, nx), dtype=vel.dtype)
    u[:, 1:-1] = 0.5 * (vel[:, :-1, 0] + vel[:, 1:, 0])
    u[:, 0] = vel[:, 0, 0]
    u[:, -1] = vel[:, -1, 0]
    v[1:-1, :] = 0.5 * (vel[:-1, :, 1] + vel[1:, :, 1])
    v[0, :] = vel[0, :, 1]
    v[-1, :] = vel[-1, :, 1]
    return u, v


def raise_sf_status(rep: N.fb_sf_report, evidence: dict | None = None) -> None:
    info = N.SF_STATUS.get(int(rep.status))
    if info is None:
        return
    reason, stage, msg = info
    ev = dict(evidence or {})
    if rep.status == 1:
        ev.update({"criterion": "update-linf", "iterations": int(rep.viscosity_iterations),
                   "final_residual": float(rep.viscosity_residual)})
    if rep.status == 8:
        ev.update({"criterion": "relative-l2", "iterations": int(rep.pressure_iterations),
                   "relative_residual": float(rep.pressure_residual)})
    raise NumericalFailure(reason, f"Stable Fluids {msg}", stage, ev)



# Pressure-solve precision of each form (fb_sf_params.mixed_poisson).
_POISSON_MODE = {"plain": 0, "mixed": 1, "precond32": 2}

class StableFluidsSolver:
    info = SolverInfo(
        id="stable-fluids",
        display_name="Stable Fluids (MAC)",
        dimensions=(2,),
        supports_moving_boundary=True,
        supported_precisions=("float32", "float64"),
        acceleration="gfx950 HIP MAC grid + DST-preconditioned CG (foilbench_amd)",
    )

    def __init__(self, form: str = "plain", ctx: N.Context | None = None, warm_pressure: int = 0) -> None:
        """`warm_pressure`: 0, each pressure CG starts from zero (the reference); 1, from the previous
        substep's pressure; 2, from its linear extrapolation 2 p_n - p_(n-1). Same stopping rule,
        fewer iterations. Not with the 'mixed' form."""
        if int(warm_pressure) not in (0, 1, 2):
            raise ValueError("warm_pressure must be 0, 1 or 2")
        if form not in _POISSON_MODE:
            raise ValueError("Stable Fluids forms: 'plain' (FP64 pressure), 'mixed' (FP32 CG + FP64 refinement) or "
                             "'precond32' (FP64 CG with the DST-I preconditioner in FP32)")
        if warm_pressure and form == "mixed":
            raise ValueError("warm_pressure needs FP64 CG: form 'plain' or 'precond32'")
        self._form = form
        self._warm = int(warm_pressure)
        self._ctx = ctx
        self._handle = None
        self._scenario: Scenario | None = None
        self._geometry: NacaFoil | None = None
        self._control = ControlState(0.0, 0.0, 0.0)
        self._time = 0.0
        self._projection_warning = ""
        self._last_projection = ("relative-l2", 0.0, 0, 0.0, True)
        self._last_viscosity = ("update-linf", 0.0, 0, 0.0, True)
        self._maccormack = True
        self._skew_rk2 = False
        self._face_advection = False
        self._reynolds = 1.0
        self._revision = 0
        self._last_refinements = 0

    @property
    def reynolds(self) -> float:
        return self._reynolds

    @property
    def state_revision(self) -> int:
        return self._revision

    def set_reynolds(self, reynolds: float) -> ReynoldsOutcome:
        if not np.isfinite(reynolds) or reynolds <= 0.0:
            raise ValueError("Reynolds number must be finite and positive")
        sel = float(reynolds)
        if sel != self._reynolds:
            self._revision += 1
        self._reynolds = sel
        return ReynoldsOutcome(self._reynolds, self._reynolds)

    @property
    def transport_mode(self) -> str:
        if self._skew_rk2:
            return "skew-rk2"
        return "maccormack" if self._maccormack else "semi-lagrangian"

    def set_transport_mode(self, mode: str) -> None:
        # fb/solvers/stable_fluids.py:169-174
        if mode not in _MODES:
            raise ValueError(f"unsupported Stable Fluids advection: {mode}")
        changed = mode != self.transport_mode
        self._maccormack = mode == "maccormack"
        self._skew_rk2 = mode == "skew-rk2"
        if changed and self._scenario is not None:
            self._revision += 1

    def interactive_tuning(self) -> InteractiveTuning:
        m = self.transport_mode
        return InteractiveTuning("stable-advection", "adv", m, m, m != "maccormack", m != "skew-rk2")

    def adjust_interactive_tuning(self, direction: int) -> InteractiveTuning:
        self.set_transport_mode("maccormack" if direction < 0 else "skew-rk2")
        return self.interactive_tuning()

    def apply_interactive_tuning(self, value) -> InteractiveTuning:
        self.set_transport_mode(str(value))
        return self.interactive_tuning()

    # ---------------------------------------------------------------- lifecycle
    def _require(self):
        if self._scenario is None or self._geometry is None or self._handle is None:
            raise RuntimeError("solver has not been initialized")
        return self._scenario, self._geometry

    def _lib(self):
        return self._ctx.lib

    def _storage_dtype(self):
        return self._scenario.dtype

    def initialize(self, scenario: Scenario, geometry: NacaFoil, seed: int) -> None:
        # fb/solvers/stable_fluids.py:195-216
        del seed
        if scenario.domain.dimension != 2:
            raise NotImplementedError("Stable Fluids supports only 2D")
        if self._ctx is None:
            self._ctx = N.Context.default()
        self._scenario = scenario
        self._geometry = geometry if isinstance(geometry, NacaFoil) else NacaFoil(geometry.spec)
        self._control = scenario.control_at(0.0)
        self._time = 0.0
        self._revision = 0
        self._reynolds = float(scenario.reynolds)
        self._last_projection = ("relative-l2", 0.0, 0, 0.0, True)
        self._last_viscosity = ("update-linf", 0.0, 0, 0.0, True)
        mode = scenario.solver_options.get("stable_advection", "maccormack")
        if mode not in _MODES:
            raise ValueError(f"unsupported Stable Fluids advection: {mode}")
        self._maccormack = mode == "maccormack"
        self._skew_rk2 = mode == "skew-rk2"
        self._face_advection = bool(scenario.solver_options.get("stable_face_advection", False))
        self._create_handle()
        u, v = cell_to_faces(initial_velocity(scenario))
        self._set_faces(u, v)
        N.check(self._lib().fb_mac_set_pose(self._handle, C.byref(pose_struct(self._control.angle_degrees))))
        self._apply_projection(max(scenario.output_dt, 1.0e-4))
        self._revision = 0

    def _create_handle(self) -> None:
        if self._handle is not None:
            self._lib().fb_mac_destroy(self._handle)
            self._handle = None
        sc = self._scenario
        prec = N.FB_F32 if sc.precision == "float32" else N.FB_F64
        h = C.c_void_p()
        N.check(self._lib().fb_mac_create(self._ctx.handle, C.byref(domain_struct(sc.domain)),
                                          C.byref(foil_struct(sc.foil)), prec, C.byref(h)))
        self._handle = h

    def _set_faces(self, u: np.ndarray, v: np.ndarray) -> None:
        dt = self._storage_dtype()
        N.check(self._lib().fb_mac_set_faces(self._handle, N.ptr(np.ascontiguousarray(u, dtype=dt)),
                                             N.ptr(np.ascontiguousarray(v, dtype=dt))))

    def faces(self) -> tuple[np.ndarray, np.ndarray]:
        sc, _ = self._require()
        dt = self._storage_dtype()
        u = np.empty((sc.domain.ny, sc.domain.nx + 1), dtype=dt)
        v = np.empty((sc.domain.ny + 1, sc.domain.nx), dtype=dt)
        N.check(self._lib().fb_mac_get_faces(self._handle, N.ptr(u), N.ptr(v)))
        return u, v

    def solid_mask(self) -> np.ndarray:
        sc, _ = self._require()
        s = np.empty((sc.domain.ny, sc.domain.nx), dtype=np.uint8)
        N.check(self._lib().fb_mac_get_geometry(self._handle, None, N.ptr(s)))
        return s.astype(bool)

    def pressure(self) -> np.ndarray:
        sc, _ = self._require()
        p = np.empty((sc.domain.ny, sc.domain.nx), dtype=np.float64)
        N.check(self._lib().fb_mac_get_pressure(self._handle, N.ptr(p)))
        return p

    def _set_pressure(self, p: np.ndarray) -> None:
        N.check(self._lib().fb_mac_set_pressure(self._handle, N.ptr(np.ascontiguousarray(p, dtype=np.float64))))

    def restart(self, scenario: Scenario, geometry: NacaFoil, seed: int, start: RestartState) -> None:
        # fb/solvers/stable_fluids.py:218-236
        validate_restart_state(start)
        self.initialize(scenario, geometry, seed)
        self.set_reynolds(start.reynolds)
        self._time = start.time
        self._control = ControlState(start.time, start.angle_degrees, 0.0)
        N.check(self._lib().fb_mac_set_pose(self._handle, C.byref(pose_struct(start.angle_degrees))))
        solid = self.solid_mask()
        vel = initial_velocity(scenario)
        wall = self._wall_grid(self._control)
        vel[solid] = wall[solid]
        u, v = cell_to_faces(vel)
        self._set_faces(u, v)
        self._apply_projection(max(scenario.output_dt, 1.0e-4))
        self._revision = 0

    def _wall_grid(self, control: ControlState) -> np.ndarray:
        sc, geometry = self._require()
        pts = cell_centers(sc.domain).reshape(-1, 2)
        return geometry.wall_velocity(pts, control).reshape(sc.domain.ny, sc.domain.nx, 2)

    def _params(self, dt: float, control: ControlState) -> N.fb_sf_params:
        sc = self._scenario
        cfl = float(sc.solver_options.get("stable_cfl", 0.7))
        p = N.fb_sf_params()
        p.dt = dt
        p.viscosity = sc.reference_speed * sc.foil.chord / self._reynolds
        p.tolerance = _float_option(sc, "pressure_tolerance", 1.0e-5)
        p.max_iterations = _int_option(sc, "pressure_max_iterations", 640)
        p.maccormack = 1 if self._maccormack else 0
        p.freestream_x = float(sc.freestream[0])
        p.freestream_y = float(sc.freestream[1])
        p.wall_omega_rad = float(np.deg2rad(control.angular_velocity_degrees))
        if self._skew_rk2:
            cfl = min(cfl, 0.4)  # fb/solvers/stable_fluids.py:265-266
        p.cfl_limit = max(_MIN_PROJECTION_CFL_LIMIT, 2.0 * cfl)
        p.channel_walls = 1 if str(sc.solver_options.get("initial_condition", "")) == "poiseuille" else 0
        p.mixed_poisson = _POISSON_MODE[self._form]
        p.warm_pressure = self._warm
        p.face_transport = 2 if self._skew_rk2 else (1 if self._face_advection else 0)
        return p

    def _record(self, rep: N.fb_sf_report, projection: bool = True, viscosity: bool = False) -> None:
        sc = self._scenario
        tol = _float_option(sc, "pressure_tolerance", 1.0e-5)
        if projection:
            self._last_projection = ("relative-l2", tol, int(rep.pressure_iterations), float(rep.pressure_residual),
                                     bool(rep.pressure_converged))
            self._last_refinements = int(rep.pressure_refinements)
        if viscosity:
            self._last_viscosity = ("update-linf", tol, int(rep.viscosity_iterations),
                                    float(rep.viscosity_residual), bool(rep.viscosity_converged))

    def _apply_projection(self, dt: float) -> None:
        # fb/solvers/stable_fluids.py:256-330
        sc, _ = self._require()
        cfl_option = sc.solver_options.get("stable_cfl", 0.7)
        if not isinstance(cfl_option, (int, float)):
            raise TypeError("stable_cfl must be numeric")
        rep = N.fb_sf_report()
        params = self._params(dt, self._control)
        N.check(self._lib().fb_mac_project(self._handle, C.byref(params), C.byref(rep)))
        if rep.status == 3:
            raise NumericalFailure("excessive_velocity",
                                   f"Stable Fluids projection CFL {rep.projection_cfl:.2f} exceeds "
                                   f"{params.cfl_limit:.2f}", "projection",
                                   {"maximum_cfl": float(rep.projection_cfl),
                                    "maximum_cfl_limit": float(params.cfl_limit)})
        if rep.status in (4, 5, 6, 7, 8, 9):
            self._record(rep)
        raise_sf_status(rep)
        self._record(rep)
        self._projection_warning = ""

    def cell_velocity(self) -> np.ndarray:
        u, v = self.faces()
        return np.stack((0.5 * (u[:, :-1] + u[:, 1:]), 0.5 * (v[:-1, :] + v[1:, :])), axis=-1)

    def _stats(self, control: ControlState) -> N.fb_mac_stats:
        st = N.fb_mac_stats()
        N.check(self._lib().fb_mac_get_stats(self._handle, float(np.deg2rad(control.angular_velocity_degrees)),
                                             C.byref(st)))
        return st

    def _skew_rate(self) -> float:
        """skew_face_advection_rate of the current faces (fb/core/grid.py:434-443), on the device."""
        r = C.c_double()
        N.check(self._lib().fb_mac_skew_rate(self._handle, C.byref(r)))
        return float(r.value)

    def advance(self, control: ControlState, target_dt: float) -> StepReport:
        # fb/solvers/stable_fluids.py:332-603
        sc, geometry = self._require()
        validate_advance_request(self._time, control, target_dt, sc.precision)
        st0 = self._stats(self._control)
        max_speed = max(st0.max_cell_speed, abs(sc.freestream[0]), 1.0e-6)
        cfl_option = sc.solver_options.get("stable_cfl", 0.7)
        if not isinstance(cfl_option, (int, float)):
            raise TypeError("stable_cfl must be numeric")
        cfl = float(cfl_option)
        if self._skew_rk2:
            cfl = min(cfl, 0.4)
        spacing = min(sc.domain.dx, sc.domain.dy)
        wall_speed = abs(np.deg2rad(control.angular_velocity_degrees)) * geometry.maximum_radius
        sweep_cells = abs(np.deg2rad(control.angle_degrees - self._control.angle_degrees)) * geometry.maximum_radius / spacing
        advective_rate = self._skew_rate() if self._skew_rk2 else max_speed / spacing
        fluid_measure = target_dt * advective_rate
        required = max(1.05 * fluid_measure / cfl, target_dt * wall_speed / (cfl * spacing), sweep_cells / cfl)
        substeps = max(1, int(np.ceil(required)))
        if substeps > 512:
            raise NumericalFailure("stability_limit", "Stable Fluids motion requires too many internal substeps",
                                   "advection" if self._skew_rk2 else "boundary",
                                   {"required_substeps": substeps, "maximum_substeps": 512,
                                                "maximum_fluid_speed": max_speed, "maximum_wall_speed": wall_speed,
                                                "boundary_sweep_cells": sweep_cells})
        dt = target_dt / substeps
        divergence_limit = _postcondition_limit(sc, "mac_maximum_divergence_linf")
        leakage_limit = _postcondition_limit(sc, "mac_maximum_solid_leakage")
        N.check(self._lib().fb_mac_checkpoint(self._handle))
        saved = (self._control, self._time, self._projection_warning, self._revision, self._last_projection,
                 self._last_viscosity)
        start_time = self._time
        start_angle = self._control.angle_degrees
        final_speed = max_speed
        accepted = fluid_measure / substeps
        retries = 0

        def restore() -> None:
            N.check(self._lib().fb_mac_rollback(self._handle))
            (self._control, self._time, self._projection_warning, self._revision, self._last_projection,
             self._last_viscosity) = saved

        while True:
            try:
                for k in range(substeps):
                    frac = (k + 1) / substeps
                    sub = ControlState(start_time + frac * target_dt,
                                       start_angle + frac * (control.angle_degrees - start_angle),
                                       control.angular_velocity_degrees)
                    rep = N.fb_sf_report()
                    params = self._params(dt, sub)
                    N.check(self._lib().fb_sf_substep(self._handle, C.byref(params),
                                                      C.byref(pose_struct(sub.angle_degrees)), C.byref(rep)))
                    self._record(rep, projection=rep.status not in (1, 2, 3), viscosity=True)
                    if rep.status == 3:
                        raise NumericalFailure("excessive_velocity",
                                               f"Stable Fluids projection CFL {rep.projection_cfl:.2f} exceeds "
                                               f"{params.cfl_limit:.2f}", "projection",
                                               {"maximum_cfl": float(rep.projection_cfl),
                                                "maximum_cfl_limit": float(params.cfl_limit)})
                    raise_sf_status(rep)
                    self._control = sub
                st = self._stats(self._control)
                if st.nonfinite_faces:
                    raise NumericalFailure("nonfinite_state", "Stable Fluids produced non-finite velocity",
                                           "postcondition")
                final_speed = st.max_cell_speed
                advective_rate = self._skew_rate() if self._skew_rk2 else final_speed / spacing
                accepted = dt * advective_rate
            except Exception:
                restore()
                raise
            if accepted <= cfl * (1.0 + 1.0e-6):
                break
            nxt = max(substeps + 1, int(np.ceil(1.05 * substeps * accepted / cfl)))
            restore()
            if nxt > 512:
                raise NumericalFailure("stability_limit", "Stable Fluids retry requires too many internal substeps",
                                       "advection", {"accepted_measure": accepted, "maximum_measure": cfl,
                                                     "required_substeps": nxt, "maximum_substeps": 512,
                                                     "stability_retries": retries})
            retries += 1
            substeps = nxt
            dt = target_dt / substeps
        final_control = ControlState(start_time + target_dt, control.angle_degrees, control.angular_velocity_degrees)
        st = self._stats(final_control)
        native_divergence = st.divergence_linf
        native_leakage = st.solid_leakage
        if native_divergence > divergence_limit or native_leakage > leakage_limit:
            restore()
            raise NumericalFailure("postcondition_failure",
                                   "Stable Fluids exceeded a configured MAC postcondition limit", "postcondition",
                                   {"divergence_linf": native_divergence, "maximum_divergence_linf": divergence_limit,
                                    "solid_leakage": native_leakage, "maximum_solid_leakage": leakage_limit})
        self._time = start_time + target_dt
        self._control = ControlState(self._time, control.angle_degrees, control.angular_velocity_degrees)
        self._revision += 1
        warnings = () if not self._projection_warning else (self._projection_warning,)
        lp = self._last_projection
        lv = self._last_viscosity
        return StepReport(target_dt, target_dt, substeps, final_speed, warnings, self._revision, {
            "maximum_fluid_speed": final_speed,
            "maximum_wall_speed": wall_speed,
            "maximum_characteristic_displacement": accepted,
            "maximum_advective_rate": advective_rate,
            "maximum_boundary_sweep": sweep_cells / substeps,
            "stability_retries": retries,
            "pressure_converged": True,
            "pressure_iterations": lp[2],
            "pressure_relative_residual": lp[3],
            "viscosity_converged": lv[4],
            "viscosity_iterations": lv[2],
            "viscosity_final_residual": lv[3],
            "divergence_linf": native_divergence,
            "solid_leakage": native_leakage,
            "requested_reynolds": self._reynolds,
            "effective_reynolds": self._reynolds,
            "degraded_motion": wall_speed == 0.0 and abs(control.angle_degrees - start_angle) > 1.0e-9,
            **({"pressure_refinements": self._last_refinements} if self._form == "mixed" else {}),
        })

    def sample_velocity(self, points: np.ndarray) -> np.ndarray:
        # fb/solvers/stable_fluids.py:609-611: the cell velocity sampled bilinearly, on the device
        sc, _ = self._require()
        pts = sample_points(points)
        out = np.empty(pts.shape, dtype=sc.dtype)
        N.check(self._lib().fb_mac_sample_velocity(self._handle, pts.shape[0], N.ptr(pts), N.ptr(out)))
        return out

    def export_state(self) -> CanonicalFlowState:
        # fb/solvers/stable_fluids.py:613-634
        sc, _ = self._require()
        vel = self.cell_velocity().copy()
        vel[self.solid_mask()] = 0.0
        return CanonicalFlowState(
            schema_version=2, dimension=2, bounds=sc.domain.bounds, resolution=sc.domain.resolution,
            periodic_axes=sc.domain.periodic_axes, time=self._time, precision=sc.precision,
            angle_degrees=self._control.angle_degrees, angular_velocity_degrees=self._control.angular_velocity_degrees,
            source_language="python", source_solver=self.info.id, velocity=vel[None, ...], geometry=sc.foil,
            producer_execution_target="native",
        )

    def import_state(self, state: CanonicalFlowState, control: ControlState) -> ImportOutcome:
        # fb/solvers/stable_fluids.py:636-706
        sc, geometry = self._require()
        N.check(self._lib().fb_mac_checkpoint(self._handle))
        saved = (self._control, self._time, self._projection_warning, self._revision, self._last_projection,
                 self._last_viscosity)

        def restore():
            N.check(self._lib().fb_mac_rollback(self._handle))
            (self._control, self._time, self._projection_warning, self._revision, self._last_projection,
             self._last_viscosity) = saved

        try:
            solid = geometry.mask(sc.domain, control.angle_degrees, self._ctx)
            validate_canonical_import(state, sc, control, solid)
            vel = np.asarray(state.velocity[0], dtype=sc.dtype).copy()
            self._time = state.time
            self._control = control
            N.check(self._lib().fb_mac_set_pose(self._handle, C.byref(pose_struct(control.angle_degrees))))
            wall = self._wall_grid(control)
            vel[solid] = wall[solid]
            u, v = cell_to_faces(vel)
            self._set_faces(u, v)
            self._apply_projection(1.0e-4)
        except NumericalFailure as failure:
            restore()
            return ImportOutcome("rejected", failure.reason, warnings=(str(failure),), stage=failure.stage,
                                 evidence=failure.evidence)
        except Exception:
            restore()
            raise
        self._revision += 1
        report = ImportReport(state.source_solver, self.info.id, ("pressure", "face-centered projection history"),
                              ("Stable Fluids rebuilt pressure and face-projection history.",))
        return ImportOutcome("accepted", "none", report, report.warnings)

    def diagnostics(self) -> Diagnostics:
        # fb/solvers/stable_fluids.py:708-738
        sc, _ = self._require()
        st = self._stats(self._control)
        d = N.fb_diag()
        N.check(self._lib().fb_mac_diagnostics(self._handle, DiagParams(sc).ref(), C.byref(d)))
        values = {
            "time": self._time,
            "requested_reynolds": self._reynolds,
            "kinetic_energy": d.kinetic_energy,
            "enstrophy": d.enstrophy,
            "divergence_l2": d.divergence_l2,
            "divergence_linf": st.divergence_linf,
            "solid_leakage": st.solid_leakage,
            "wake_width": d.wake_width,
            "recirculation_area": d.recirculation_area,
        }
        if not all(np.isfinite(v) for v in values.values()):
            raise NumericalFailure("nonfinite_state", "Stable Fluids produced non-finite diagnostics")
        warnings = () if not self._projection_warning else (self._projection_warning,)
        return Diagnostics(values, warnings, self._revision)

    def forces(self) -> dict:
        """Force on the foil per unit depth (unit density) with drag, lift and their coefficients."""
        sc, _ = self._require()
        f = (C.c_double * 2)()
        nu = sc.reference_speed * sc.foil.chord / self._reynolds
        N.check(self._lib().fb_mac_forces(self._handle, nu, float(np.deg2rad(self._control.angular_velocity_degrees)),
                                          f))
        return lift_drag(f[0], f[1], sc.freestream[:2], sc.foil.chord)

    def close(self) -> None:
        """Free the device state now rather than at garbage collection."""
        if self._handle is not None and self._ctx is not None:
            self._ctx.lib.fb_mac_destroy(self._handle)
            self._handle = None

    def __del__(self) -> None:  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass
