"""Value types of the solver lifecycle: domain, foil, controls, scenario, reports, failures.

These mirror the reference's shared model objects so that a caller written against
``foilbench_py.core.models`` (implementations/python/src/foilbench_py/core/models.py:11-391)
finds the same names, fields, validation rules and error types here.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Literal

import numpy as np

Precision = Literal["float32", "float64"]
AxisName = Literal["x", "y", "z"]

# Reason and stage vocabularies of classified failures (models.py:14-44).
FAILURE_STAGES = (
    "canonical-import", "advection", "viscosity", "projection", "boundary", "collision",
    "streaming", "particle-transfer", "particle-advection", "population-maintenance",
    "time-mapping", "postcondition",
)
FAILURE_REASONS = (
    "none", "excessive_velocity", "stability_limit", "nonfinite_state", "convergence_failure",
    "incompatible_geometry", "incompatible_domain", "projection_failure", "invalid_density",
    "invalid_population", "invalid_relaxation", "transfer_failure", "postcondition_failure",
    "time_contract_failure", "unsupported_conversion",
)


def _finite(*values: float) -> bool:
    return all(math.isfinite(float(v)) for v in values)


@dataclass(frozen=True, slots=True)
class DomainSpec:
    """Axis-aligned box with a uniform cell grid (models.py:50-93)."""

    dimension: Literal[2, 3]
    bounds: tuple[tuple[float, float], ...]
    resolution: tuple[int, ...]
    periodic_axes: tuple[AxisName, ...] = ()

    def __post_init__(self) -> None:
        if len(self.bounds) != self.dimension or len(self.resolution) != self.dimension:
            raise ValueError("bounds and resolution must match dimension")
        for lower, upper in self.bounds:
            if not _finite(lower, upper) or upper <= lower:
                raise ValueError("each domain upper bound must exceed its lower bound")
        if any(n < 4 for n in self.resolution):
            raise ValueError("each resolution axis must contain at least four cells")
        allowed = ("x", "y") if self.dimension == 2 else ("x", "y", "z")
        if len(set(self.periodic_axes)) != len(self.periodic_axes) or any(
            a not in allowed for a in self.periodic_axes
        ):
            raise ValueError("periodic axes must be unique and valid for the dimension")

    @property
    def nx(self) -> int:
        return self.resolution[0]

    @property
    def ny(self) -> int:
        return self.resolution[1]

    @property
    def nz(self) -> int:
        return 1 if self.dimension == 2 else self.resolution[2]

    @property
    def dx(self) -> float:
        return (self.bounds[0][1] - self.bounds[0][0]) / self.nx

    @property
    def dy(self) -> float:
        return (self.bounds[1][1] - self.bounds[1][0]) / self.ny


@dataclass(frozen=True, slots=True)
class FoilSpec:
    """NACA four-digit foil, chord length and pivot (models.py:96-108)."""

    naca: str
    chord: float
    pivot: tuple[float, ...]

    def __post_init__(self) -> None:
        if len(self.naca) != 4 or not self.naca.isdigit():
            raise ValueError("NACA code must contain exactly four digits")
        if not math.isfinite(self.chord) or self.chord <= 0.0:
            raise ValueError("foil chord must be positive")
        if not self.pivot or not _finite(*self.pivot):
            raise ValueError("foil pivot must contain finite coordinates")


@dataclass(frozen=True, slots=True)
class ControlKeyframe:
    time: float
    angle_degrees: float

    def __post_init__(self) -> None:
        if not math.isfinite(self.time) or self.time < 0.0:
            raise ValueError("control keyframe time must be finite and nonnegative")
        if not math.isfinite(self.angle_degrees) or not -90.0 <= self.angle_degrees <= 90.0:
            raise ValueError("control keyframe angle must be finite and within [-90, 90]")


@dataclass(frozen=True, slots=True)
class ControlState:
    time: float
    angle_degrees: float
    angular_velocity_degrees: float

    def __post_init__(self) -> None:
        if not _finite(self.time, self.angle_degrees, self.angular_velocity_degrees):
            raise ValueError("control state must be finite")
        if self.time < 0.0:
            raise ValueError("control state time cannot be negative")


@dataclass(frozen=True, slots=True)
class Scenario:
    """A validated scenario (models.py:143-224)."""

    schema_version: int
    id: str
    domain: DomainSpec
    reynolds: float
    freestream: tuple[float, ...]
    foil: FoilSpec
    controls: tuple[ControlKeyframe, ...]
    duration: float
    output_dt: float
    precision: Precision
    seed: int
    solver_options: dict[str, object] = field(default_factory=dict)

    def __post_init__(self) -> None:
        if self.schema_version != 1 or not self.id:
            raise ValueError("scenario schema version and id are required")
        if self.domain.dimension != len(self.freestream):
            raise ValueError("freestream dimensionality must match the domain")
        if len(self.foil.pivot) != self.domain.dimension:
            raise ValueError("foil pivot dimensionality must match the domain")
        if not self.controls:
            raise ValueError("a scenario needs at least one control keyframe")
        if any(b.time <= a.time for a, b in zip(self.controls, self.controls[1:])):
            raise ValueError("control keyframe times must be strictly increasing")
        if not _finite(self.reynolds, self.duration, self.output_dt, *self.freestream):
            raise ValueError("scenario physical values must be finite")
        if self.reynolds <= 0.0 or self.duration <= 0.0 or self.output_dt <= 0.0:
            raise ValueError("Reynolds number, duration, and output_dt must be positive")
        initial = str(self.solver_options.get("initial_condition", "freestream"))
        periodic = set(self.domain.periodic_axes)
        if initial == "taylor-green" and (self.domain.dimension != 2 or periodic != {"x", "y"}):
            raise ValueError("Taylor-Green requires a 2D domain periodic in x and y")
        if initial == "poiseuille" and (
            self.domain.dimension != 2 or "x" not in periodic or "y" in periodic
        ):
            raise ValueError("Poiseuille requires a 2D domain periodic in x and nonperiodic in y")

    @property
    def dtype(self) -> np.dtype:
        return np.dtype(np.float32 if self.precision == "float32" else np.float64)

    @property
    def reference_speed(self) -> float:
        speed = float(np.linalg.norm(np.asarray(self.freestream, dtype=np.float64)))
        initial = str(self.solver_options.get("initial_condition", "freestream"))
        prescribed = 1.0 if initial in ("taylor-green", "poiseuille") else 0.0
        return max(speed, prescribed, 1.0e-12)

    def control_at(self, time: float) -> ControlState:
        """Smoothstep angle schedule and its derivative (models.py:203-224)."""
        keys = self.controls
        if len(keys) == 1 or time <= keys[0].time:
            return ControlState(time, keys[0].angle_degrees, 0.0)
        if time >= keys[-1].time:
            return ControlState(time, keys[-1].angle_degrees, 0.0)
        for left, right in zip(keys, keys[1:]):
            if left.time <= time <= right.time:
                span = right.time - left.time
                if span <= 0.0:
                    return ControlState(time, right.angle_degrees, 0.0)
                s = (time - left.time) / span
                smooth = s * s * (3.0 - 2.0 * s)
                angle = left.angle_degrees + smooth * (right.angle_degrees - left.angle_degrees)
                rate = 6.0 * s * (1.0 - s) * (right.angle_degrees - left.angle_degrees) / span
                return ControlState(time, angle, rate)
        return ControlState(time, keys[-1].angle_degrees, 0.0)


@dataclass(frozen=True, slots=True)
class SolverInfo:
    id: str
    display_name: str
    dimensions: tuple[int, ...]
    supports_moving_boundary: bool
    supported_precisions: tuple[str, ...]
    acceleration: str


@dataclass(frozen=True, slots=True)
class InteractiveTuning:
    id: str
    label: str
    value: str | float
    display_value: str
    can_decrease: bool
    can_increase: bool


@dataclass(frozen=True, slots=True)
class StepReport:
    requested_dt: float
    advanced_dt: float
    substeps: int
    max_speed: float
    warnings: tuple[str, ...] = ()
    state_revision: int = 0
    evidence: dict[str, object] = field(default_factory=dict)


@dataclass(frozen=True, slots=True)
class RestartState:
    time: float
    angle_degrees: float
    reynolds: float


@dataclass(frozen=True, slots=True)
class ReynoldsOutcome:
    requested: float
    effective: float
    warnings: tuple[str, ...] = ()


@dataclass(frozen=True, slots=True)
class ImportReport:
    source_solver: str
    destination_solver: str
    discarded_state: tuple[str, ...]
    warnings: tuple[str, ...] = ()


@dataclass(frozen=True, slots=True)
class ImportOutcome:
    status: Literal["accepted", "rejected"]
    reason: str
    report: ImportReport | None = None
    warnings: tuple[str, ...] = ()
    stage: str | None = None
    evidence: dict[str, object] = field(default_factory=dict)

    @property
    def accepted(self) -> bool:
        return self.status == "accepted"


class NumericalFailure(RuntimeError):
    """Classified, anticipated numerical failure; the solver state is rolled back first."""

    def __init__(
        self,
        reason: str,
        detail: str,
        stage: str = "postcondition",
        evidence: dict[str, object] | None = None,
    ) -> None:
        super().__init__(detail)
        self.reason = reason
        self.stage = stage
        self.evidence = {} if evidence is None else evidence


@dataclass(frozen=True, slots=True)
class Diagnostics:
    values: dict[str, float]
    warnings: tuple[str, ...] = ()
    state_revision: int = 0


@dataclass(frozen=True, slots=True)
class CanonicalFlowState:
    """Language-neutral exported flow state (models.py:321-391)."""

    schema_version: int
    dimension: Literal[2, 3]
    bounds: tuple[tuple[float, float], ...]
    resolution: tuple[int, ...]
    periodic_axes: tuple[AxisName, ...]
    time: float
    precision: Precision
    angle_degrees: float
    angular_velocity_degrees: float
    source_language: str
    source_solver: str
    velocity: np.ndarray
    density: np.ndarray | None = None
    geometry: FoilSpec | None = None
    producer_execution_target: str | None = None

    def __post_init__(self) -> None:
        if self.schema_version not in (1, 2):
            raise ValueError("unsupported canonical schema version")
        if self.schema_version == 2 and (self.geometry is None or not self.producer_execution_target):
            raise ValueError("canonical v2 requires geometry and producer target identity")
        if len(self.bounds) != self.dimension or len(self.resolution) != self.dimension:
            raise ValueError("canonical bounds and resolution must match dimension")
        for lower, upper in self.bounds:
            if not _finite(lower, upper) or upper <= lower:
                raise ValueError("canonical bounds must be finite and increasing")
        if any(n < 1 for n in self.resolution):
            raise ValueError("canonical resolution axes must be positive")
        allowed = ("x", "y") if self.dimension == 2 else ("x", "y", "z")
        if len(set(self.periodic_axes)) != len(self.periodic_axes) or any(
            a not in allowed for a in self.periodic_axes
        ):
            raise ValueError("canonical periodic axes are invalid for the dimension")
        if not _finite(self.time, self.angle_degrees, self.angular_velocity_degrees):
            raise ValueError("canonical time and control values must be finite")
        if self.time < 0.0:
            raise ValueError("canonical time cannot be negative")
        if not self.source_language or not self.source_solver:
            raise ValueError("canonical source language and solver are required")
        nz = 1 if self.dimension == 2 else self.resolution[2]
        shape = (nz, self.resolution[1], self.resolution[0], self.dimension)
        if self.velocity.shape != shape:
            raise ValueError(f"velocity shape {self.velocity.shape} does not match {shape}")
        if self.density is not None and self.density.shape != shape[:-1]:
            raise ValueError("density shape does not match the canonical grid")
        want = np.dtype(self.precision)
        if self.velocity.dtype != want:
            raise TypeError("velocity dtype does not match canonical precision")
        if self.density is not None and self.density.dtype != want:
            raise TypeError("density dtype does not match canonical precision")
        if not np.isfinite(self.velocity).all():
            raise ValueError("canonical velocity contains non-finite values")
        if self.density is not None and not np.isfinite(self.density).all():
            raise ValueError("canonical density contains non-finite values")
