"""Serving leg: the request mix served by the native request loop against the plain form.

Three runs over the same mix (PCG32 seed 42, stream 1001; step counts scaled by `step_scale` so that the
leg finishes in about a minute):
* capacity: every request queued at t = 0, served with continuous batching (Stable Fluids in FP64 storage
  with the drop-in's FP64 CG, the reference's zero start, same-shape pressure solves batched; lattices in
  shifted FP32 storage with FP64 collision, batched launches) -> aggregate cell-steps/s;
* plain: the drop-in FP64 form, one request after another -> aggregate cell-steps/s;
* latency: Poisson arrivals at the mix's rate on the wall clock -> p50/p99 completion latency, and the
  aggregate rate under those arrivals.
Launch shapes and the lattice chunk are picked by the autotuner (`foilbench_amd.tuning`) for the mix's grid
shapes before the runs. The relative L2 field error of the served results against the plain FP64 results is
measured on the first requests of each solver in the scaled mix, and at the mix's own horizons on the first
request of each solver and grid shape with its full step count (200-5000 steps): velocity faces for Stable
Fluids; density deviation and velocity for the lattice.
"""

from __future__ import annotations

import numpy as np

from ..serving.engine import Server, run_plain
from ..serving.mix import LBM, SF, serving_mix


def _rel(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _field_errors(served: dict, plain_outcomes, only=None) -> dict:
    errors = {}
    for o in plain_outcomes:
        if (only is not None and o.id not in only) or o.id not in served or not served[o.id].fields or not o.fields:
            continue
        s = served[o.id]
        for k, ref in o.fields.items():
            if k == "rho":
                errors[f"{o.solver}:{o.id}:rho-1"] = _rel(s.fields[k] - 1.0, ref - 1.0)
            else:
                errors[f"{o.solver}:{o.id}:{k}"] = _rel(s.fields[k], ref)
    return errors


def full_horizon_errors(n: int = 256, rate: float = 50.0, sf_workers: int = 8) -> dict:
    """Served against plain FP64 at full step counts: the first request of each solver and grid shape."""
    mix = serving_mix(n, rate=rate)
    first = {}
    for r in mix:
        first.setdefault((r.solver, r.nx, r.ny), r)
    reqs = [first[k] for k in sorted(first)]
    srv = Server(sf_workers=sf_workers)
    try:
        served = {o.id: o for o in srv.run(reqs, keep_fields=True, realtime=False)["outcomes"]}
    finally:
        srv.close()
    plain = run_plain(reqs, keep_fields=True)
    errors = _field_errors(served, plain["outcomes"])
    return {"requests": [{"id": r.id, "solver": r.solver, "grid": f"{r.nx}x{r.ny}", "steps": r.steps} for r in reqs],
            "max_rel_field_error": max(errors.values()) if errors else None, "field_errors": errors}


def serving_leg(n: int = 256, step_scale: float = 0.05, rate: float = 50.0, sf_workers: int = 16,
                lbm_chunk: int = 32, check_per_solver: int = 4, autotune: bool = True) -> dict:
    mix = serving_mix(n, rate=rate, step_scale=step_scale)
    check = set()
    for solver in (SF, LBM):
        check.update([r.id for r in mix if r.solver == solver][:check_per_solver])
    srv = Server.tuned(mix, sf_workers=sf_workers) if autotune else Server(sf_workers=sf_workers, lbm_chunk=lbm_chunk)
    try:
        cap = srv.run(mix, keep_fields=check, realtime=False)
        lat = srv.run(mix, realtime=True)
    finally:
        srv.close()
    plain = run_plain(mix, keep_fields=check)
    errors = _field_errors({o.id: o for o in cap["outcomes"]}, plain["outcomes"], check)
    full = full_horizon_errors(n, rate)
    strip = lambda d: {k: v for k, v in d.items() if k != "outcomes"}  # noqa: E731
    ratio = cap["cell_steps_per_s"] / plain["cell_steps_per_s"]
    return {
        "metric": "aggregate cell-steps/s, serving form vs plain form",
        "value": cap["cell_steps_per_s"],
        "unit": "cell-steps/s",
        "plain_value": plain["cell_steps_per_s"],
        "ratio_vs_plain": ratio,
        "ratio_under_arrivals": lat["cell_steps_per_s"] / plain["cell_steps_per_s"],
        "p50_ms": lat["p50_ms"],
        "p99_ms": lat["p99_ms"],
        "max_rel_field_error": max([*errors.values(), full["max_rel_field_error"] or 0.0]) if errors else None,
        "field_errors": errors,
        "full_horizon": full,
        "workload": (f"serving mix: {n} requests, Poisson arrivals {rate:g}/s, grids 128x64..1024x512, "
                     f"SF/LBM 50/50, steps 200..5000 x {step_scale:g}"),
        "tuning": srv.tuning,
        "capacity_run": strip(cap),
        "latency_run": strip(lat),
        "plain_run": strip(plain),
    }


def full_mix_leg(n: int = 256, rate: float = 50.0, sf_workers: int = 16, chunk: int = 16) -> dict:
    """The mix at its own horizons (200-5000 steps per request, step scale 1): the capacity run of the
    serving form, then the plain form over the same requests (in chunks, with a line printed after each),
    and the relative field error of every served request against its plain FP64 result."""
    import time

    mix = serving_mix(n, rate=rate, step_scale=1.0)
    ids = {r.id for r in mix}
    srv = Server.tuned(mix, sf_workers=sf_workers)
    try:
        cap = srv.run(mix, keep_fields=ids, realtime=False)
    finally:
        srv.close()
    print(f"capacity: {cap['cell_steps_per_s']:.4g} cell-steps/s over {cap['makespan_s']:.1f} s", flush=True)
    served = {o.id: o for o in cap["outcomes"]}
    order = sorted(mix, key=lambda q: (q.arrival_s, q.id))
    cells = seconds = 0.0
    errors = {}
    for k in range(0, len(order), chunk):
        part = order[k:k + chunk]
        t0 = time.perf_counter()
        plain = run_plain(part, keep_fields=True)
        seconds += time.perf_counter() - t0
        cells += float(sum(o.cells * o.steps for o in plain["outcomes"] if o.status == 0 and o.sf_status == 0))
        errors.update(_field_errors(served, plain["outcomes"]))
        print(f"plain: {k + len(part)}/{len(order)} requests, {cells / seconds:.4g} cell-steps/s", flush=True)
    by_solver = {}
    for key, v in errors.items():
        s = key.split(":")[0]
        by_solver[s] = max(by_solver.get(s, 0.0), v)
    strip = lambda d: {k: v for k, v in d.items() if k != "outcomes"}  # noqa: E731
    return {
        "workload": f"serving mix: {n} requests, grids 128x64..1024x512, SF/LBM 50/50, steps 200..5000 (full)",
        "value": cap["cell_steps_per_s"],
        "plain_value": cells / seconds,
        "ratio_vs_plain": cap["cell_steps_per_s"] / (cells / seconds),
        "max_rel_field_error": max(errors.values()) if errors else None,
        "max_rel_field_error_by_solver": by_solver,
        "fields_checked": len(errors),
        "capacity_run": strip(cap),
        "plain_seconds": seconds,
    }


if __name__ == "__main__":
    import json
    import sys

    if len(sys.argv) > 1 and sys.argv[1] == "full":
        print(json.dumps(full_mix_leg(), default=str), flush=True)
    else:
        print(json.dumps(serving_leg(), default=str), flush=True)
