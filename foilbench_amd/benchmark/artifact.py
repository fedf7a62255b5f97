"""Result-v2 artifact rules (fb/benchmark/artifact.py:1-176): the numerical configuration a result records,
its physical identity, and the cross-field checks the JSON schema cannot express.

These are the interchange contract of the reference's benchmark artifacts, restated so that the GPU runner's
results validate exactly as the reference's do and sit beside them in its comparer.
"""

from __future__ import annotations

import math
from collections.abc import Mapping, Sequence

# fb/benchmark/artifact.py:9-21: the solver options a result records, with their defaults (None: recorded
# only when the scenario sets it)
SOLVER_OPTION_DEFAULTS: dict[str, object] = {
    "initial_condition": "freestream",
    "stable_advection": "maccormack",
    "stable_face_advection": False,
    "stable_cfl": 0.7,
    "pressure_tolerance": 1.0e-5,
    "pressure_max_iterations": 640,
    "mac_maximum_divergence_linf": None,
    "mac_maximum_solid_leakage": None,
    "pic_flip_blend": 0.95,
    "pic_population_interval": 8,
    "pic_cfl": 0.75,
}

PHYSICAL_IDENTITY_FIELDS = ("bounds", "periodic_axes", "reynolds", "effective_reynolds", "solver_configuration",
                            "freestream", "foil", "control_history", "requested_duration", "output_dt", "seed")


def solver_configuration(scenario) -> dict:
    """fb/benchmark/artifact.py:24-31."""
    opts = scenario.solver_options
    return {k: opts.get(k, d) for k, d in SOLVER_OPTION_DEFAULTS.items() if d is not None or k in opts}


def physical_identity(result: Mapping) -> dict:
    """fb/benchmark/artifact.py:34-50."""
    return {k: result[k] for k in PHYSICAL_IDENTITY_FIELDS}


def identities_match(a, b, *, precision: str) -> bool:
    """fb/benchmark/artifact.py:53-97: numbers within 2e-6 (float32) or 2e-12 (float64), relative and absolute;
    booleans, integers and strings exactly; mappings and sequences element by element."""
    if isinstance(a, bool) or isinstance(b, bool):
        return type(a) is type(b) and a == b
    if isinstance(a, int) and isinstance(b, int):
        return a == b
    if isinstance(a, (int, float)) and isinstance(b, (int, float)):
        tol = 2.0e-6 if precision == "float32" else 2.0e-12
        return math.isclose(float(a), float(b), rel_tol=tol, abs_tol=tol)
    if isinstance(a, Mapping) and isinstance(b, Mapping):
        return a.keys() == b.keys() and all(identities_match(a[k], b[k], precision=precision) for k in a)
    seq = lambda v: isinstance(v, Sequence) and not isinstance(v, (str, bytes, bytearray))  # noqa: E731
    if seq(a) and seq(b):
        return len(a) == len(b) and all(identities_match(x, y, precision=precision) for x, y in zip(a, b))
    if isinstance(a, (str, bytes, bytearray)) or isinstance(b, (str, bytes, bytearray)):
        return type(a) is type(b) and a == b
    return a is None and b is None


def _finite_everywhere(value, where: str = "result") -> None:
    if isinstance(value, float):
        if not math.isfinite(value):
            raise ValueError(f"{where} contains a non-finite number")
    elif isinstance(value, Mapping):
        for k, v in value.items():
            _finite_everywhere(v, f"{where}.{k}")
    elif isinstance(value, Sequence) and not isinstance(value, (str, bytes, bytearray)):
        for i, v in enumerate(value):
            _finite_everywhere(v, f"{where}[{i}]")


def _linear_percentile(ordered: list[float], fraction: float) -> float:
    pos = fraction * (len(ordered) - 1)
    lo, hi = math.floor(pos), math.ceil(pos)
    return ordered[lo] + (ordered[hi] - ordered[lo]) * (pos - lo)


def validate_result_semantics(result: Mapping) -> None:
    """fb/benchmark/artifact.py:117-176: finite numbers throughout; a successful result has no failure, a last
    step and diagnostics of the final revision, and covered its requested duration (1e-6 relative in float32,
    1e-12 in float64); a failed one carries structured failure evidence; the derived rates agree with the step
    times (1e-10 relative); the memory kind agrees with the RSS value."""
    _finite_everywhere(result)
    steps = [float(v) for v in result["step_seconds"]]
    requested = float(result["requested_duration"])
    simulated = float(result["simulated_duration"])
    tol = (1.0e-6 if result["precision"] == "float32" else 1.0e-12) * max(1.0, abs(requested))
    revision = result["final_state_revision"]
    if result["success"]:
        last = result["last_step"]
        if result["failure"] is not None or not isinstance(last, Mapping) or not steps:
            raise ValueError("successful benchmark result lacks completed-step semantics")
        if result["diagnostic_state_revision"] != revision or last.get("state_revision") != revision:
            raise ValueError("successful benchmark result contains stale revision evidence")
        if abs(simulated - requested) > tol:
            raise ValueError("successful benchmark result did not complete requested duration")
    elif not isinstance(result["failure"], Mapping):
        raise ValueError("failed benchmark result lacks structured failure evidence")
    if steps:
        ordered = sorted(steps)
        wall = sum(steps)
        cells = math.prod(int(v) for v in result["resolution"])
        subs = int(result["substeps"])
        particles = float(result["diagnostics"].get("particle_count", 0.0))
        derived = {
            "median_step_seconds": _linear_percentile(ordered, 0.5),
            "p95_step_seconds": _linear_percentile(ordered, 0.95),
            "simulated_seconds_per_wall_second": simulated / wall,
            "cell_updates_per_second": cells * subs / wall,
            "particle_updates_per_second": particles * subs / wall,
        }
        for k, want in derived.items():
            if not math.isclose(float(result[k]), want, rel_tol=1.0e-10, abs_tol=1.0e-12):
                raise ValueError(f"benchmark result contains inconsistent derived field {k}")
    if (result["memory_measurement"] == "unavailable") != (result["peak_rss_bytes"] is None):
        raise ValueError("memory measurement kind and RSS value disagree")
