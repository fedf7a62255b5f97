"""Scenario documents (fb/core/scenario.py:35-91): JSON validated against the scenario schema,
then built into the immutable ``Scenario`` model."""

from __future__ import annotations

import json
from pathlib import Path

from .models import ControlKeyframe, DomainSpec, FoilSpec, Scenario
from .schema import load_schema, validate


def _load_object(path: Path) -> dict:
    obj = json.loads(path.read_text(encoding="utf-8"))
    if not isinstance(obj, dict):
        raise TypeError(f"{path} must contain a JSON object")
    return obj


def load_scenario(path: str | Path) -> Scenario:
    return load_scenario_document(_load_object(Path(path)))


def load_scenario_document(raw: dict, schema: dict | None = None) -> Scenario:
    """Validate an in-memory scenario document and construct the model."""
    validate(raw, schema if schema is not None else load_schema("scenario.schema.json"))
    domain = DomainSpec(
        dimension=int(raw["dimension"]),
        bounds=tuple((float(p[0]), float(p[1])) for p in raw["bounds"]),
        resolution=tuple(int(v) for v in raw["resolution"]),
        periodic_axes=tuple(str(v) for v in raw["periodic_axes"]),
    )
    f = raw["foil"]
    foil = FoilSpec(naca=str(f["naca"]), chord=float(f["chord"]), pivot=tuple(float(v) for v in f["pivot"]))
    controls = tuple(ControlKeyframe(time=float(c["time"]), angle_degrees=float(c["angle_degrees"]))
                     for c in raw["controls"])
    return Scenario(
        schema_version=int(raw["schema_version"]),
        id=str(raw["id"]),
        domain=domain,
        reynolds=float(raw["reynolds"]),
        freestream=tuple(float(v) for v in raw["freestream"]),
        foil=foil,
        controls=controls,
        duration=float(raw["duration"]),
        output_dt=float(raw["output_dt"]),
        precision=str(raw["precision"]),
        seed=int(raw["seed"]),
        solver_options=dict(raw.get("solver_options", {})),
    )


def scenario_document(sc: Scenario) -> dict:
    """The JSON document of a scenario (inverse of load_scenario_document)."""
    doc = {
        "schema_version": sc.schema_version,
        "id": sc.id,
        "dimension": sc.domain.dimension,
        "bounds": [list(b) for b in sc.domain.bounds],
        "resolution": list(sc.domain.resolution),
        "periodic_axes": list(sc.domain.periodic_axes),
        "reynolds": sc.reynolds,
        "freestream": list(sc.freestream),
        "foil": {"naca": sc.foil.naca, "chord": sc.foil.chord, "pivot": list(sc.foil.pivot)},
        "controls": [{"time": c.time, "angle_degrees": c.angle_degrees} for c in sc.controls],
        "duration": sc.duration,
        "output_dt": sc.output_dt,
        "precision": sc.precision,
        "seed": sc.seed,
    }
    if sc.solver_options:
        doc["solver_options"] = dict(sc.solver_options)
    return doc
