"""Benchmark-matrix runner writing result-v2 artifacts (fb/benchmark/runner.py:182-482), on the GPU solvers.

For every resolution x solver x repetition of a benchmark matrix (spec/schemas/benchmark-matrix.schema.json):
a cold solver is initialised and advanced once (both timed, excluded from the rates), a fresh solver is
initialised and advanced output step by output step over the matrix duration with each ``advance`` timed by
``perf_counter``; the wake probe (1.5 chords behind the pivot) is sampled over the second half of the run,
the wake recovery after an angle excursion is tracked, and the run's diagnostics close it. The result document
is validated against result-v2.schema.json and by ``validate_result_semantics``, written as JSON with the
canonical state beside it when the matrix asks for snapshots, and summarised in summary.csv.

Result documents stay schema-valid for the reference's comparer: ``implementation`` is "python" and
``execution_target`` "native" (the schema's vocabulary); the GPU backend and the solver form are recorded in the
free-form ``machine`` object (fb/benchmark/artifact.py:34-50 keeps ``solver_configuration`` physical).

    python -m foilbench_amd.benchmark.runner MATRIX.json [--root DIR] [--out DIR] [--form plain]
"""

from __future__ import annotations

import argparse
import csv
import json
import os
import platform
import statistics
import subprocess
import sys
import time
from dataclasses import dataclass, replace
from pathlib import Path

import numpy as np

from ..core.geometry import NacaFoil
from ..core.metrics import analyze_wake_probe
from ..core.models import NumericalFailure
from ..core.scenario import load_scenario
from ..core.schema import load_schema, validate
from ..core.state_io import save_canonical_state
from ..solvers.factory import create_solver
from .artifact import solver_configuration, validate_result_semantics

RECOVERY_OBSERVATION_LIMIT = 4.0
RECOVERY_TIME_TOLERANCE = 1.0e-9
RECOVERY_EVENT_TOLERANCE = 1.0e-12


@dataclass(frozen=True)
class Matrix:
    id: str
    scenario_path: Path
    solvers: tuple
    resolutions: tuple
    duration: float
    repetitions: int
    save_snapshots: bool


def load_matrix(path: str | Path, root: str | Path | None = None) -> Matrix:
    """fb/benchmark/runner.py:63-78: a schema-valid matrix whose scenario path is relative to `root` (by default
    the nearest parent directory of the matrix that holds that path)."""
    path = Path(path).resolve()
    raw = json.loads(path.read_text(encoding="utf-8"))
    validate(raw, load_schema("benchmark-matrix.schema.json"))
    if root is None:
        root = next((p for p in path.parents if (p / raw["scenario"]).exists()), path.parent)
    return Matrix(id=str(raw["id"]), scenario_path=Path(root) / raw["scenario"], solvers=tuple(raw["solvers"]),
                  resolutions=tuple((int(r[0]), int(r[1])) for r in raw["resolutions"]),
                  duration=float(raw["duration"]), repetitions=int(raw["repetitions"]),
                  save_snapshots=bool(raw["save_snapshots"]))


def recovery_window(sc) -> tuple[float, float] | None:
    """fb/benchmark/runner.py:155-179: the end of the baseline before an angle excursion and the time the foil is
    back at its initial angle, when the schedule returns to it before the run ends."""
    first, last = sc.controls[0].angle_degrees, sc.controls[-1].angle_degrees
    if not np.isclose(first, last):
        return None
    moved = [i for i, k in enumerate(sc.controls) if not np.isclose(k.angle_degrees, first)]
    if not moved or moved[0] == 0 or moved[-1] + 1 >= len(sc.controls):
        return None
    start, end = sc.controls[moved[0] - 1].time, sc.controls[moved[-1] + 1].time
    if start >= end or end >= sc.duration:
        return None
    return start, end


def aligned_step(elapsed: float, dt: float, events) -> float:
    """fb/benchmark/runner.py:32-42: a step shortened so it lands on a scheduled landmark."""
    if events is not None:
        for t in events:
            until = t - elapsed
            if RECOVERY_EVENT_TOLERANCE < until < dt:
                dt = until
    return dt


def _scalar(v):
    if isinstance(v, np.generic):
        v = v.item()
    if isinstance(v, float) and not np.isfinite(v):
        raise ValueError("benchmark evidence values must be finite")
    if v is None or isinstance(v, (bool, int, float, str)):
        return v
    raise TypeError(f"benchmark evidence value {v!r} is not a JSON scalar")


def _step_record(rep):
    """fb/benchmark/runner.py:125-136."""
    if rep is None:
        return None
    return {"requested_dt": rep.requested_dt, "advanced_dt": rep.advanced_dt, "substeps": rep.substeps,
            "max_speed": rep.max_speed, "state_revision": rep.state_revision,
            "evidence": {k: _scalar(v) for k, v in rep.evidence.items()}, "warnings": list(rep.warnings)}


def _failure_record(err: Exception) -> dict:
    """fb/benchmark/runner.py:139-152."""
    if isinstance(err, NumericalFailure):
        return {"kind": "numerical", "reason": err.reason, "stage": err.stage, "message": str(err),
                "evidence": {k: _scalar(v) for k, v in err.evidence.items()}}
    return {"kind": "unexpected", "reason": None, "stage": None, "message": f"{type(err).__name__}: {err}",
            "evidence": {}}


def _git_commit(root: Path) -> str:
    try:
        return subprocess.run(["git", "rev-parse", "HEAD"], cwd=root, check=True, capture_output=True,
                              text=True).stdout.strip() or "unknown"
    except (OSError, subprocess.CalledProcessError):
        return "unknown"


def _machine(form: str) -> dict:
    from .. import _native as N

    out = {"platform": platform.platform(), "python": platform.python_version(), "processor": platform.processor(),
           "logical_cpus": os.cpu_count(), "numpy": np.__version__, "backend": "foilbench_amd (HIP, gfx950)",
           "solver_form": form}
    try:
        n = N.Context.default().device
        out["device"] = n
    except Exception:
        pass
    return out


def _percentile95(v: list[float]) -> float:
    return v[0] if len(v) == 1 else float(np.percentile(np.asarray(v), 95.0))


def run_case(sc, solver_id: str, *, matrix_id: str, duration: float, repetition: int, form: str = "plain",
             git_commit: str = "unknown") -> tuple[dict, object]:
    """One result-v2 document (fb/benchmark/runner.py:208-450) and the solver it ran (None on failure)."""
    import psutil

    proc = psutil.Process()
    geometry = NacaFoil(sc.foil)
    init_s = cold_s = 0.0
    solver = None
    elapsed = 0.0
    step_s: list[float] = []
    substeps = 0
    peak = proc.memory_info().rss
    warnings: list[str] = []
    probe: list[float] = []
    events = recovery_window(sc)
    base = None
    recovered = None
    last = None
    diag_rev = None
    failure = None
    diag: dict = {}
    probe_pt = np.asarray([[min(sc.foil.pivot[0] + 1.5 * sc.foil.chord, sc.domain.bounds[0][1] - 0.5 * sc.domain.dx),
                            sc.foil.pivot[1]]], dtype=sc.dtype)
    ok = True
    try:
        cold = create_solver(solver_id, form)
        t0 = time.perf_counter()
        cold.initialize(sc, geometry, sc.seed)
        init_s = time.perf_counter() - t0
        wdt = min(sc.output_dt, sc.duration)
        t0 = time.perf_counter()
        cold.advance(sc.control_at(wdt), wdt)
        cold_s = time.perf_counter() - t0
        del cold
        solver = create_solver(solver_id, form)
        solver.initialize(sc, geometry, sc.seed)
        while elapsed < sc.duration - 1.0e-12:
            if events is not None and base is None and elapsed >= events[0] - RECOVERY_EVENT_TOLERANCE:
                v = solver.diagnostics().values
                base = (v["wake_width"], v["recirculation_area"])
            dt = aligned_step(elapsed, min(sc.output_dt, sc.duration - elapsed), events)
            control = sc.control_at(elapsed + dt)
            t0 = time.perf_counter()
            rep = solver.advance(control, dt)
            step_s.append(time.perf_counter() - t0)
            last = rep
            elapsed += rep.advanced_dt
            substeps += rep.substeps
            warnings.extend(rep.warnings)
            peak = max(peak, proc.memory_info().rss)
            if elapsed >= 0.5 * sc.duration:
                probe.append(float(solver.sample_velocity(probe_pt)[0, 1]))
            if events is not None:
                crossed = base is None and elapsed >= events[0] - RECOVERY_EVENT_TOLERANCE
                watching = base is not None and recovered is None and elapsed >= events[1] - RECOVERY_EVENT_TOLERANCE
                if crossed or watching:
                    v = solver.diagnostics().values
                    wake, recirc = v["wake_width"], v["recirculation_area"]
                    if crossed:
                        base = (wake, recirc)
                    elif (wake <= max(1.25 * base[0], 2.0 * sc.domain.dy)
                          and recirc <= max(1.25 * base[1], 2.0 * sc.domain.dx * sc.domain.dy)):
                        recovered = elapsed - events[1]
        d = solver.diagnostics()
        if d.state_revision != solver.state_revision:
            raise RuntimeError("benchmark diagnostics describe a stale state revision")
        diag_rev = d.state_revision
        warnings.extend(d.warnings)
        diag = dict(d.values)
        if len(probe) >= 8:
            spec = analyze_wake_probe(np.asarray(probe, dtype=np.float64), sc.output_dt, sc.foil.chord,
                                      max(float(np.linalg.norm(sc.freestream)), 1.0e-12))
            diag.update({"wake_probe_samples": float(spec.sample_count),
                         "wake_frequency_resolution": spec.frequency_resolution,
                         "wake_transverse_rms": spec.transverse_rms,
                         "wake_mixing_index": spec.transverse_rms / sc.reference_speed,
                         "wake_dominant_frequency": spec.dominant_frequency,
                         "wake_strouhal_number": spec.strouhal_number,
                         "wake_dominant_power_fraction": spec.dominant_power_fraction})
        if events is not None and base is not None:
            limit = min(RECOVERY_OBSERVATION_LIMIT, max(0.0, sc.duration - events[1]))
            seen = recovered is not None and recovered <= limit + RECOVERY_TIME_TOLERANCE
            diag.update({"recovery_baseline_time": events[0], "recovery_start_time": events[1],
                         "recovery_observed": float(seen),
                         "recovery_elapsed": min(recovered, limit) if seen else limit})
            if not seen:
                warnings.append("wake recovery was not observed; recovery_elapsed is right-censored")
    except Exception as err:  # a failed run is an artifact too (runner.py:377-382)
        ok = False
        warnings.append(f"{type(err).__name__}: {err}")
        diag = {}
        diag_rev = None
        failure = _failure_record(err)
    wall = sum(step_s)
    cells = sc.domain.nx * sc.domain.ny
    particles = diag.get("particle_count", 0.0)
    result = {
        "schema_version": 2, "contract_id": "foilbench-phase3-v1", "contract_revision": 5,
        "benchmark_matrix_id": matrix_id, "scenario_id": sc.id, "repetition": repetition,
        "language": "python", "implementation": "python", "execution_target": "native", "solver": solver_id,
        "git_commit": git_commit, "machine": _machine(form), "precision": sc.precision,
        "resolution": list(sc.domain.resolution), "bounds": [list(b) for b in sc.domain.bounds],
        "periodic_axes": list(sc.domain.periodic_axes), "reynolds": sc.reynolds,
        "effective_reynolds": diag.get("effective_reynolds", sc.reynolds),
        "solver_configuration": solver_configuration(sc), "freestream": list(sc.freestream),
        "foil": {"naca": sc.foil.naca, "chord": sc.foil.chord, "pivot": list(sc.foil.pivot)},
        "control_history": [{"time": k.time, "angle_degrees": k.angle_degrees} for k in sc.controls],
        "requested_duration": duration, "simulated_duration": elapsed, "output_dt": sc.output_dt, "seed": sc.seed,
        "initialization_seconds": init_s, "cold_step_seconds": cold_s, "step_seconds": step_s,
        "median_step_seconds": statistics.median(step_s) if step_s else 0.0,
        "p95_step_seconds": _percentile95(step_s) if step_s else 0.0,
        "simulated_seconds_per_wall_second": elapsed / wall if wall > 0.0 else 0.0,
        "cell_updates_per_second": cells * substeps / wall if wall > 0.0 else 0.0,
        "particle_updates_per_second": particles * substeps / wall if wall > 0.0 else 0.0,
        "peak_rss_bytes": peak, "memory_measurement": "rss", "runtime_startup_seconds": None,
        "worker_startup_seconds": None, "substeps": substeps,
        "final_state_revision": 0 if solver is None else solver.state_revision,
        "diagnostic_state_revision": diag_rev, "last_step": _step_record(last), "diagnostics": diag,
        "success": ok, "failure": failure, "warnings": sorted(set(warnings)),
    }
    validate(result, load_schema("result-v2.schema.json"))
    validate_result_semantics(result)
    return result, (solver if ok else None)


def run_matrix(matrix_path: str | Path, output_root: str | Path | None = None, *, root: str | Path | None = None,
               form: str = "plain") -> Path:
    """fb/benchmark/runner.py:182-482 on the GPU solvers; returns the output directory."""
    m = load_matrix(matrix_path, root)
    out = Path(output_root) if output_root else Path("results") / m.id / time.strftime("%Y%m%d-%H%M%S")
    out.mkdir(parents=True, exist_ok=True)
    base = load_scenario(m.scenario_path)
    commit = _git_commit(Path(__file__).resolve().parents[2])
    rows = []
    for res in m.resolutions:
        sc = replace(base, domain=replace(base.domain, resolution=res), duration=m.duration)
        for sid in m.solvers:
            for rep in range(m.repetitions):
                result, solver = run_case(sc, sid, matrix_id=m.id, duration=m.duration, repetition=rep + 1,
                                          form=form, git_commit=commit)
                stem = f"{sid}-{res[0]}x{res[1]}-r{rep + 1}"
                (out / f"{stem}.json").write_text(json.dumps(result, indent=2, allow_nan=False), encoding="utf-8")
                if m.save_snapshots and result["success"] and solver is not None:
                    save_canonical_state(solver.export_state(), out / f"{stem}-state")
                rows.append({"solver": sid, "resolution": f"{res[0]}x{res[1]}", "repetition": rep + 1,
                             **{k: result[k] for k in ("initialization_seconds", "cold_step_seconds",
                                                       "median_step_seconds", "p95_step_seconds",
                                                       "simulated_seconds_per_wall_second",
                                                       "cell_updates_per_second", "particle_updates_per_second",
                                                       "peak_rss_bytes", "success")}})
    if rows:
        with (out / "summary.csv").open("w", newline="", encoding="utf-8") as fh:
            w = csv.DictWriter(fh, fieldnames=list(rows[0]))
            w.writeheader()
            w.writerows(rows)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("matrix")
    ap.add_argument("--root", default=None, help="directory the matrix's scenario path is relative to")
    ap.add_argument("--out", default=None)
    ap.add_argument("--form", default="plain")
    a = ap.parse_args()
    print(run_matrix(a.matrix, a.out, root=a.root, form=a.form), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
