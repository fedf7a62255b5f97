"""Canonical flow-state directories (fb/core/state_io.py:48-187): ``manifest.json`` plus
little-endian C-order ``velocity.npy`` (and ``density.npy``), manifest v1 or v2. Arrays are
written and read with allow_pickle=False."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np

from .models import CanonicalFlowState, FoilSpec
from .schema import SchemaError, load_schema, validate


def save_canonical_state(state: CanonicalFlowState, directory: str | Path) -> Path:
    out = Path(directory)
    out.mkdir(parents=True, exist_ok=True)
    dt = np.dtype(state.precision).newbyteorder("<")
    np.save(out / "velocity.npy", np.ascontiguousarray(np.asarray(state.velocity, dtype=state.precision).astype(dt)),
            allow_pickle=False)
    if state.density is not None:
        np.save(out / "density.npy",
                np.ascontiguousarray(np.asarray(state.density, dtype=state.precision).astype(dt)), allow_pickle=False)
    m: dict[str, object] = {
        "schema_version": state.schema_version,
        "dimension": state.dimension,
        "bounds": [list(b) for b in state.bounds],
        "resolution": list(state.resolution),
        "periodic_axes": list(state.periodic_axes),
        "time": state.time,
        "precision": state.precision,
        "angle_degrees": state.angle_degrees,
        "angular_velocity_degrees": state.angular_velocity_degrees,
        "source_solver": state.source_solver,
        "velocity": {"file": "velocity.npy", "axes": ["z", "y", "x", "component"], "order": "C"},
        "density": None if state.density is None else {"file": "density.npy", "axes": ["z", "y", "x"],
                                                      "order": "C"},
    }
    if state.schema_version == 1:
        m["source_language"] = state.source_language
    else:
        if state.geometry is None or state.producer_execution_target is None:
            raise ValueError("canonical v2 requires geometry and producer target")
        g = state.geometry
        m["geometry"] = {"family": "naca-four-digit-v1", "naca": g.naca, "chord": g.chord, "pivot": list(g.pivot)}
        m["producer"] = {"implementation": state.source_language, "execution_target": state.producer_execution_target}
    (out / "manifest.json").write_text(json.dumps(m, indent=2), encoding="utf-8")
    return out


def _array_meta(m: dict, name: str, axes: list[str]) -> dict | None:
    raw = m.get(name)
    if raw is None:
        return None
    if not isinstance(raw, dict):
        raise TypeError(f"canonical {name} metadata must be an object")
    if raw.get("file") != f"{name}.npy":
        raise ValueError(f"canonical {name} must use {name}.npy")
    if raw.get("axes") != axes:
        raise ValueError(f"canonical {name} axes must be {axes}")
    if raw.get("order") not in ("C", "F"):
        raise ValueError(f"canonical {name} must declare C or Fortran order")
    return raw


def _order_ok(a: np.ndarray, meta: dict) -> bool:
    return bool(a.flags.c_contiguous if meta["order"] == "C" else a.flags.f_contiguous)


def load_canonical_state(directory: str | Path) -> CanonicalFlowState:
    src = Path(directory)
    m = json.loads((src / "manifest.json").read_text(encoding="utf-8"))
    if not isinstance(m, dict):
        raise TypeError("canonical manifest must be a JSON object")
    version = int(m.get("schema_version", 0)) if isinstance(m.get("schema_version", 0), int) else 0
    schema = load_schema("canonical-manifest.schema.json" if version == 1 else "canonical-manifest-v2.schema.json")
    try:
        validate(m, schema)
    except SchemaError as err:
        raise ValueError(f"invalid canonical manifest at {err.path or '$'}: {err.message}") from err
    vmeta = _array_meta(m, "velocity", ["z", "y", "x", "component"])
    if vmeta is None:
        raise ValueError("canonical velocity metadata is required")
    dmeta = _array_meta(m, "density", ["z", "y", "x"])
    precision = str(m["precision"])
    if precision not in ("float32", "float64"):
        raise ValueError("canonical precision must be float32 or float64")
    want = np.dtype(precision)
    vel = np.load(src / "velocity.npy", allow_pickle=False)
    if vel.dtype != want or not _order_ok(vel, vmeta):
        raise ValueError("canonical velocity dtype/order does not match its manifest")
    dens = None
    if dmeta is not None:
        dens = np.load(src / "density.npy", allow_pickle=False)
        if dens.dtype != want or not _order_ok(dens, dmeta):
            raise ValueError("canonical density dtype/order does not match its manifest")
    dim = int(m["dimension"])
    if dim not in (2, 3):
        raise ValueError("canonical dimension must be 2 or 3")
    geom = None
    if isinstance(m.get("geometry"), dict):
        g = m["geometry"]
        geom = FoilSpec(naca=str(g["naca"]), chord=float(g["chord"]), pivot=tuple(float(v) for v in g["pivot"]))
    lang = str(m.get("source_language", ""))
    target = None
    if isinstance(m.get("producer"), dict):
        lang = str(m["producer"]["implementation"])
        target = str(m["producer"]["execution_target"])
    return CanonicalFlowState(
        schema_version=int(m["schema_version"]),
        dimension=dim,
        bounds=tuple((float(p[0]), float(p[1])) for p in m["bounds"]),
        resolution=tuple(int(v) for v in m["resolution"]),
        periodic_axes=tuple(str(v) for v in m["periodic_axes"]),
        time=float(m["time"]),
        precision=precision,
        angle_degrees=float(m["angle_degrees"]),
        angular_velocity_degrees=float(m["angular_velocity_degrees"]),
        source_language=lang,
        source_solver=str(m["source_solver"]),
        velocity=np.asarray(vel, dtype=want),
        density=None if dens is None else np.asarray(dens, dtype=want),
        geometry=geom,
        producer_execution_target=target,
    )


def midspan_velocity(state: CanonicalFlowState) -> np.ndarray:
    return state.velocity[state.velocity.shape[0] // 2]
