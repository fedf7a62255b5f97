"""Headline benchmark: cell-steps/s of the D2Q9 TRT LBM lattice step on one MI355X.

Workload (one "step" = one advance of the whole lattice by LATTICE_STEPS lattice steps, as the solver's advance
runs them: the last writes the pre-collision state with the postcondition reductions; inputs resident in HBM):
NACA 2412 at AoA 12 degrees (angle_degrees = -12), 1024 x 512 lattice, chord 160 cells,
equilibrium inflow u = 0.05 lattice units, Re = 1000, TRT magic parameter 3/16,
interpolated bounce-back on the foil, FP32 population storage.

Prints one JSON line. Besides the headline FP32 value it reports the FP64 plain form and the
shifted FP32/FP16 precision forms (throughput and relative velocity error against FP64),
the roofline of the step kernel from HIP events, and the oracle timed on the host's cores (the C restatement of
the lattice step on one pinned core and on all cores, with the CPU model).
Under "legs" it measures the other paths through their solver lifecycles, each with a model
roofline and an oracle baseline on the host:
  sf_plain   Stable Fluids FP64, NACA 0012 AoA 8, 256 x 128, Re 500, CG to 1e-6;
  sf_mixed   Stable Fluids mixed (FP32 CG sweeps, FP64 refinement), NACA 0012 AoA 15,
             2048 x 1024, Re 2000, CG to 1e-8, with its error against the FP64 plain form;
  sf_precond32  the sf_mixed workload with FP64 CG and the DST-I preconditioner sweeps in FP32
             (the serving form's pressure solve), with its error against the FP64 plain form;
  pic        PIC/FLIP FP32, NACA 4412 AoA 10, 1024 x 512, 8 particles per fluid cell.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-scale S] [--no-legs]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
ROOFLINE_LAUNCHES = 4000  # launches in the event-timed sample of the step kernel
# lattice steps per bench step: the reference runner times whole advances (fb/benchmark/runner.py:240-275), and an
# advance of 100 steps keeps the host's submission and synchronisation (~80 us per timed window) out of the rate
LATTICE_STEPS = 100


FORMS = (("f32", "FB_F32"), ("f64", "FB_F64"), ("f32-mixed", "FB_F32_MIXED"), ("f16-mixed", "FB_F16_MIXED"))
ERROR_STEPS = 20000  # lattice steps of the per-form error against FP64: the LBM configuration's horizon
SERVING_ERROR_STEPS = 5000  # the serving mix's longest request
SERVING_SHAPES = ((128, 64), (256, 128), (512, 256), (1024, 512))
LARGE_SHAPE = (4096, 2048)  # working set above the 256 MB Infinity Cache (604 MB in FP32)
LARGE_LAUNCHES = 1000


def _time_form(tuner, case, name, pname, ctx, steps, warmup, launches):
    """Throughput over `steps` back-to-back advances of LATTICE_STEPS lattice steps and the step kernel's
    event-timed average over `launches` launches, with the autotuner's launch shape for this lattice and storage
    form."""
    from foilbench_amd import _native as N
    from foilbench_amd.benchmark.lbm_harness import BYTES_PER_CELL, Lattice

    prec = getattr(N, pname)
    # fused pairs, interior rows per thread and the special-cell blocks are the autotuner's picks for this
    # lattice and storage form (none of them changes a result bit)
    rec = tuner.lbm(case.nx, case.ny, name)
    tuned = {k: rec[k] for k in ("fuse", "rows", "list_blocks", "nt")}
    lat = Lattice(case, prec, ctx)
    try:
        for _ in range(warmup):
            lat.run(LATTICE_STEPS)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            lat.run(LATTICE_STEPS)
        ctx.synchronize()
        wall = time.perf_counter() - t0
        # average launch duration: one HIP event pair around `launches` back-to-back launches on the kernel's
        # stream, whatever --steps is
        avg_ms = lat.time(launches) / launches
    finally:
        lat.close()
    bytes_launch = BYTES_PER_CELL[prec] * lat.cells
    per_launch = 2 if tuned["fuse"] else 1  # lattice steps per launch of the step kernel
    return {
        "cell_steps_per_s": lat.cells * LATTICE_STEPS * steps / wall,
        "ms_per_step": 1e3 * wall / steps,
        "lattice_steps_per_step": LATTICE_STEPS,
        "kernel_ms": avg_ms,
        "kernel_gbs": bytes_launch / (avg_ms * 1e-3) / 1e9,
        "bytes_per_cell": BYTES_PER_CELL[prec],
        "tuned_step_shape": tuned,
        "kernel": (f"lbm_pair_kernel<Prec<{pname}>, 16>" if tuned["fuse"]
                   else f"lbm_step_kernel<Prec<{pname}>, {tuned['rows']}, {'true' if tuned['nt'] == 1 else 'false'}, true>"),
        "steps_per_launch": per_launch,
        "launch_ms": avg_ms * per_launch,
        "event_launches": launches,
    }


def _gpu_forms(steps: int, warmup: int):
    from foilbench_amd import _native as N
    from foilbench_amd.benchmark.lbm_harness import LatticeCase
    from foilbench_amd.tuning import Autotuner

    case = LatticeCase()
    ctx = N.Context.default()
    tuner = Autotuner(ctx)
    out = {}
    for name, pname in FORMS:
        out[name] = _time_form(tuner, case, name, pname, ctx, steps, warmup, ROOFLINE_LAUNCHES)
        print(f"[bench] {name}: {out[name]['cell_steps_per_s']:.4g} cell-steps/s, kernel "
              f"{out[name]['kernel_ms'] * 1e3:.2f} us ({out[name]['kernel_gbs']:.0f} GB/s)", file=sys.stderr, flush=True)
    return case, out, tuner


def _form_errors(nx: int, ny: int, steps: int, forms=("f32", "f32-mixed", "f16-mixed")) -> dict:
    """Relative L2 velocity error over the fluid cells of each storage form against FP64 after exactly
    `steps` lattice steps from the same inflow equilibrium (NACA 2412 at AoA 12, u = 0.05, Re 1000)."""
    import numpy as np

    from foilbench_amd import _native as N
    from foilbench_amd.benchmark.lbm_harness import Lattice, LatticeCase

    ctx = N.Context.default()
    case = LatticeCase(nx=nx, ny=ny)
    vel = {}
    solid = None
    for name in ("f64",) + tuple(forms):
        lat = Lattice(case, getattr(N, dict(FORMS)[name]), ctx)
        try:
            lat.run(steps)
            ctx.synchronize()
            vel[name] = lat.velocity().astype(np.float64)
            if solid is None:
                solid = lat.solid()
        finally:
            lat.close()
    fl = ~solid
    ref = vel["f64"][fl]
    return {name: float(np.linalg.norm(vel[name][fl] - ref) / np.linalg.norm(ref)) for name in forms}


def _serving_lattice_errors() -> dict:
    """The serving form's lattices (shifted FP32 storage, FP64 collision) against the plain FP64 form after
    SERVING_ERROR_STEPS steps, for one request of each grid shape of the serving mix (NACA 2412 at AoA 12,
    Re 1000; the lattice viscosity clamp of the solver lifecycle applies at the small shapes): relative L2
    velocity error over the fluid cells. Each lattice is stepped on its own stream; the request loop's batched
    launches compute the same bits (tests/test_gpu_serving.py)."""
    import ctypes as C

    import numpy as np

    from foilbench_amd import _native as N
    from foilbench_amd.serving.engine import _build, _close
    from foilbench_amd.serving.mix import LBM, Request

    ctx = N.Context.default()
    out = {}
    for nx, ny in SERVING_SHAPES:
        req = Request(0, 0.0, LBM, nx, ny, "2412", 12.0, 1000.0, SERVING_ERROR_STEPS)
        vel = {}
        for form in ("plain", "serving"):
            s, handle, params, _ = _build(req, form, ctx)
            try:
                N.check(ctx.lib.fb_lbm_run_fixed(handle, C.byref(params), req.steps))
                ctx.synchronize()
                _, v = s._macroscopic()
                vel[form] = np.asarray(v, np.float64)
                solid = s.geometry_fields()[1]
            finally:
                _close(s)
        fl = ~solid
        ref = vel["plain"][fl]
        out[f"served_lattice_{nx}x{ny}_{SERVING_ERROR_STEPS}_steps"] = float(
            np.linalg.norm(vel["serving"][fl] - ref) / np.linalg.norm(ref))
    return out


def _large_lattice(tuner) -> dict:
    """The step kernel on a 4096 x 2048 lattice (chord 640 cells), whose populations (604 MB in FP32) cannot
    stay in the 256 MB Infinity Cache between steps: the HBM-resident figure."""
    from foilbench_amd import _native as N
    from foilbench_amd.benchmark.lbm_harness import LatticeCase

    ctx = N.Context.default()
    case = LatticeCase(nx=LARGE_SHAPE[0], ny=LARGE_SHAPE[1])
    out = {"workload": case.name, "working_set_mb": {}}
    for name, pname in FORMS:
        if name == "f64":
            continue
        r = _time_form(tuner, case, name, pname, ctx, 2, 1, LARGE_LAUNCHES)
        r["frac"] = r["kernel_gbs"] / HBM_PEAK_GBS
        out[name] = r
        out["working_set_mb"][name] = r["bytes_per_cell"] * case.nx * case.ny / 1e6  # two population buffers
        print(f"[bench] large {name}: kernel {r['kernel_ms'] * 1e3:.1f} us ({r['kernel_gbs']:.0f} GB/s)",
              file=sys.stderr, flush=True)
    return out


CPU_STEPS = {"lbm": 80, "sf_plain": 50, "sf_mixed": 1, "pic": 1}


def _cpu_run(leg_args: list[str], threads: int, pin: int | None) -> dict:
    """One oracle baseline (oracle/baseline.py) in a child process whose OpenMP and BLAS thread counts are set
    before NumPy loads, pinned to one core when `pin` is given."""
    import subprocess

    env = dict(os.environ)
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        env[k] = str(threads)
    cmd = [sys.executable, "-m", "oracle.baseline", *leg_args, "--threads", str(threads)]
    if pin is not None:
        cmd += ["--pin", str(pin)]
    res = subprocess.run(cmd, cwd=str(ROOT), env=env, capture_output=True, text=True, timeout=600)
    if res.returncode != 0:
        return {"error": res.stderr[-2000:]}
    out = json.loads(res.stdout.strip().splitlines()[-1])
    out["kind"] = "port"
    return out


def _cpu_baselines(scale: float) -> tuple[dict, dict]:
    """The headline lattice's baseline (the C restatement of the lattice step on one pinned core, and on all
    the cores this process may use) and the legs' baselines (the NumPy/SciPy oracle on one pinned core)."""
    from oracle.baseline import cpu_model

    cpus = sorted(os.sched_getaffinity(0))
    pin = cpus[0]
    all_cores = min(len(cpus), 16)
    host = {"nproc": os.cpu_count(), "affinity_cpus": len(cpus), "cpu_model": cpu_model()}
    n = lambda k: max(1, int(round(CPU_STEPS[k] * scale)))  # noqa: E731
    lbm1 = _cpu_run(["lbm", "--steps", str(n("lbm"))], 1, pin)
    lbmn = _cpu_run(["lbm", "--steps", str(4 * n("lbm"))], all_cores, None)
    head = dict(lbm1)
    head["host"] = host
    head["all_cores"] = {k: lbmn.get(k) for k in ("value", "unit", "cores", "sample", "median_advance_ms")}
    legs = {
        "sf_plain": _cpu_run(["sf", "--nx", "256", "--ny", "128", "--angle", "-8", "--re", "500", "--tol", "1e-6",
                              "--steps", str(n("sf_plain"))], 1, pin),
        # the 2048 x 1024 workload at a quarter of its cells (one oracle advance at full size takes ~40 s):
        # the rate per cell-substep of the NumPy/SciPy oracle is nearly size-independent (CG iterations grow
        # from 20 to 22 between the two sizes)
        "sf_mixed": _cpu_run(["sf", "--nx", "1024", "--ny", "512", "--angle", "-15", "--re", "2000", "--tol", "1e-8",
                              "--steps", str(n("sf_mixed"))], 1, pin),
        # the 1024 x 512 workload at 256 x 128 (one oracle advance at full size takes about a minute)
        "pic": _cpu_run(["pic", "--nx", "256", "--ny", "128", "--per-cell", "8", "--steps", str(n("pic"))], 1, pin),
    }
    for v in legs.values():
        v["host"] = host
    return head, legs


def _legs(cpu_legs: dict | None) -> dict:
    """Stable Fluids and PIC/FLIP workloads; a failing leg is reported, not fatal. `cpu_legs`: the oracle
    baselines of the legs (sf_precond32 shares sf_mixed's FP64 oracle workload)."""
    from foilbench_amd.benchmark import legs as L
    from foilbench_amd.benchmark.serving import serving_leg

    plan = {
        "sf_plain": lambda: L.stable_fluids_leg(256, 128, -8.0, 500.0, "plain", 1e-6, steps=30),
        "sf_mixed": lambda: L.stable_fluids_leg(2048, 1024, -15.0, 2000.0, "mixed", 1e-8, steps=4, warmup=1,
                                                reference_form="plain"),
        "sf_precond32": lambda: L.stable_fluids_leg(2048, 1024, -15.0, 2000.0, "precond32", 1e-8, steps=4,
                                                    warmup=1, reference_form="plain"),
        "pic": lambda: L.pic_leg(1024, 512, 8, steps=4),
        "serving": lambda: serving_leg(),
    }
    cpu_of = {"sf_plain": "sf_plain", "sf_mixed": "sf_mixed", "sf_precond32": "sf_mixed", "pic": "pic"}
    out = {}
    for name, gpu in plan.items():
        t0 = time.perf_counter()
        try:
            out[name] = gpu()
            if cpu_legs and name in cpu_of:
                out[name]["cpu_baseline"] = cpu_legs[cpu_of[name]]
        except Exception as err:  # reported in the JSON line instead of aborting the headline
            out[name] = {"error": f"{type(err).__name__}: {err}"}
        print(f"[bench] leg {name}: {time.perf_counter() - t0:.1f} s "
              f"{out[name].get('value', out[name].get('error'))}", file=sys.stderr, flush=True)
    return out


def _traffic(kernel: str):
    """HBM bytes per launch of `kernel`, from the committed rocprofv3 --pmc summary when it was
    measured on that kernel (otherwise None)."""
    p = ROOT / "profiles" / "lbm_step_kernel_1024x512_f32_traffic.json"
    if not p.exists():
        return None
    try:
        rec = json.loads(p.read_text())
        return float(rec["bytes_per_launch"]) if rec.get("kernel") == kernel else None
    except Exception:
        return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="advances of LATTICE_STEPS lattice steps (200: the "
                    "LBM configuration's 20,000-step horizon)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-scale", type=float, default=1.0,
                    help="scale of the CPU baselines' step counts (0: no CPU baseline)")
    ap.add_argument("--no-legs", action="store_true", help="measure only the LBM headline")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the path does not shard "
              "across GPUs in this build)", file=sys.stderr)
        return 2
    case, forms, tuner = _gpu_forms(args.steps, args.warmup)
    errors = {"lattice_1024x512_20000_steps": _form_errors(case.nx, case.ny, ERROR_STEPS)}
    errors.update(_serving_lattice_errors())
    for name in forms:
        if name != "f64":
            forms[name]["velocity_rel_err_vs_f64"] = errors["lattice_1024x512_20000_steps"][name]
    print(f"[bench] errors vs FP64: {errors}", file=sys.stderr, flush=True)
    large = _large_lattice(tuner)
    main_form = forms["f32"]
    achieved = main_form["kernel_gbs"]
    traffic = _traffic(main_form["kernel"])
    cpu, cpu_legs = _cpu_baselines(args.cpu_scale) if args.cpu_scale > 0 else (None, None)
    legs = None if args.no_legs else _legs(cpu_legs)
    line = {
        "metric": "cell-steps/s/GPU",
        "value": main_form["cell_steps_per_s"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_form["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": case.name + " fp32 storage, fp64 kernel arithmetic, interpolated bounce-back",
            "lattice": [case.nx, case.ny],
            "cells": case.nx * case.ny,
            "lattice_steps_per_step": LATTICE_STEPS,
            "parallelism": "single-gpu",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": main_form["kernel"],
            "steps_per_launch": main_form["steps_per_launch"],
            "launch_ms": main_form["launch_ms"],
            "bytes_per_cell": main_form["bytes_per_cell"],
        },
        "cpu_baseline": cpu,
        "forms": forms,
        "velocity_rel_err_vs_f64": errors,
        "hbm_resident": large,
        "legs": legs,
    }
    sv = (legs or {}).get("serving") or {}
    if "value" in sv:
        line["serving"] = {k: sv[k] for k in ("metric", "value", "unit", "plain_value", "ratio_vs_plain",
                                              "ratio_under_arrivals", "p50_ms", "p99_ms", "max_rel_field_error",
                                              "workload")}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
