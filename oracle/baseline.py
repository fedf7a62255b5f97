"""CPU baselines of the benchmark workloads on the host cores (test infrastructure: the oracle timed as the
reported baseline, never the product).

Run by bench.py's cpu_baseline leg as a child process, so that the thread counts of OpenMP and the BLAS are
fixed before NumPy loads and the process can be pinned to one core:

    python -m oracle.baseline lbm --steps 100 --threads 1 --pin 0
    python -m oracle.baseline sf --nx 256 --ny 128 --angle -8 --re 500 --tol 1e-6 --steps 50 --pin 0
    python -m oracle.baseline pic --nx 256 --ny 128 --per-cell 8 --steps 1 --pin 0

Each prints one JSON object. The timing follows the reference runner (fb/benchmark/runner.py:240-275,
381-385): the cold initialize and the first advance are excluded, each advance is timed with perf_counter, and
the rate is nx*ny*substeps / (sum of the timed advances), with the median and p95 of the advance times. The
lattice leg restates the reference's numba kernels as C loops (oracle/lbm_oracle.c, FP64 arithmetic, storage
rounding), the Stable Fluids and PIC/FLIP legs are the NumPy/SciPy oracle (oracle/mac.py, oracle/pic.py).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _stats(times, cells_substeps, unit="cell-steps/s"):
    import numpy as np

    t = np.asarray(times)
    return {"value": cells_substeps / float(t.sum()), "unit": unit, "median_advance_ms": 1e3 * float(np.median(t)),
            "p95_advance_ms": 1e3 * float(np.percentile(t, 95)), "seconds": float(t.sum())}


def lbm(steps: int, threads: int) -> dict:
    """The headline lattice (NACA 2412 at AoA 12, 1024 x 512, u = 0.05, Re 1000, FP32 storage): `steps` lattice
    steps of the C restatement after one warm-up step, each step timed."""
    import numpy as np

    from foilbench_amd.benchmark.lbm_harness import LatticeCase

    from . import lbm as L
    from . import lbm_c
    from .geometry import NacaFoil, cell_centers

    case = LatticeCase()
    nx, ny = case.nx, case.ny
    dx = (case.bounds[0][1] - case.bounds[0][0]) / nx
    dy = (case.bounds[1][1] - case.bounds[1][0]) / ny
    foil = NacaFoil(case.naca, case.chord, case.pivot)
    sdf = foil.signed_distance(cell_centers(case.bounds, nx, ny).reshape(-1, 2), case.angle_degrees).reshape(ny, nx)
    solid = sdf <= 0.0
    u = case.u_lattice
    vel = np.zeros((ny, nx, 2), np.float32)
    vel[..., 0] = u
    f = L.equilibrium(np.ones((ny, nx), np.float32), vel)
    feqb = f[0, 0].copy()
    outlet = f[:, -1:].copy()
    sp = L.sponge_strength_profile(nx, ny, False, False, False)
    _, op, om = case.relaxation()
    args = (solid, sdf, op, om, u, (np.float32(u), np.float32(0.0)), feqb, sp)
    tail = (case.bounds, dx, dy, case.pivot, 0.0, u, False, False, False)
    f, outlet = lbm_c.kernel_steps(f, *args, outlet, *tail, 1, threads=threads)
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        f, outlet = lbm_c.kernel_steps(f, *args, outlet, *tail, 1, threads=threads)
        times.append(time.perf_counter() - t0)
    out = _stats(times, nx * ny * steps)
    out["sample"] = (f"C restatement of the lattice step (oracle/lbm_oracle.c, -O2 -ffp-contract=off), same "
                     f"1024x512 FP32 lattice, {steps} lattice steps after 1 warm-up step, {out['seconds']:.1f} s, "
                     f"{threads} OpenMP thread(s)")
    return out


def sf(nx: int, ny: int, angle: float, re: float, tol: float, steps: int) -> dict:
    """Stable Fluids FP64 (NACA 0012, unit chord, [-1.5, 2.5] x [-1, 1], output_dt = dx/2, stable_cfl 1)."""
    from foilbench_amd.benchmark.legs import sf_scenario

    from .mac import StableFluidsOracle

    sc = sf_scenario(nx, ny, angle, re, "float64", tol)
    o = StableFluidsOracle()
    o.initialize(sc)
    t = sc.output_dt
    o.advance(sc.control_at(t), sc.output_dt)
    times, subs = [], 0
    for _ in range(steps):
        t += sc.output_dt
        t0 = time.perf_counter()
        r = o.advance(sc.control_at(t), sc.output_dt)
        times.append(time.perf_counter() - t0)
        subs += r["substeps"]
    out = _stats(times, nx * ny * subs)
    out["sample"] = (f"oracle/mac.py StableFluidsOracle (NumPy/SciPy cg with the DST-I preconditioner), "
                     f"{nx}x{ny}, NACA0012 AoA{-angle:g}, Re{re:g}, tol {tol:g}, {steps} advances after 1 warm-up, "
                     f"{out['seconds']:.1f} s")
    return out


def pic(nx: int, ny: int, per_cell: int, steps: int) -> dict:
    """PIC/FLIP FP32 (NACA 4412 at AoA 10 on [-1.6, 4.8] x [-1.6, 1.6], blend 0.95, PCG32 seed 42)."""
    from foilbench_amd.benchmark.legs import pic_scenario

    from .pic import PicFlipOracle

    sc = pic_scenario(nx, ny, "float32")
    o = PicFlipOracle()
    o.initialize(sc, seed=42, per_cell=per_cell)
    t = sc.output_dt
    o.advance(sc.control_at(t), sc.output_dt)
    times, subs = [], 0
    for _ in range(steps):
        t += sc.output_dt
        t0 = time.perf_counter()
        r = o.advance(sc.control_at(t), sc.output_dt)
        times.append(time.perf_counter() - t0)
        subs += r["substeps"]
    n = int(o.positions.shape[0])
    out = _stats(times, nx * ny * subs)
    out["particle_steps_per_s"] = n * subs / out["seconds"]
    out["sample"] = (f"oracle/pic.py PicFlipOracle (NumPy), {nx}x{ny}, {per_cell}/cell ({n} particles), {steps} "
                     f"advances after 1 warm-up, {out['seconds']:.1f} s")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("leg", choices=("lbm", "sf", "pic"))
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--threads", type=int, default=1)
    ap.add_argument("--pin", type=int, default=-1, help="pin the process to this CPU")
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ny", type=int, default=128)
    ap.add_argument("--angle", type=float, default=-8.0)
    ap.add_argument("--re", type=float, default=500.0)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--per-cell", type=int, default=8)
    a = ap.parse_args()
    if a.pin >= 0:
        os.sched_setaffinity(0, {a.pin})
    if a.leg == "lbm":
        out = lbm(a.steps, a.threads)
    elif a.leg == "sf":
        out = sf(a.nx, a.ny, a.angle, a.re, a.tol, a.steps)
    else:
        out = pic(a.nx, a.ny, a.per_cell, a.steps)
    out["cores"] = a.threads if a.leg == "lbm" else 1
    out["pinned_cpu"] = a.pin if a.pin >= 0 else None
    out["affinity_cpus"] = len(os.sched_getaffinity(0))
    out["threads_env"] = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS",
                                                         "MKL_NUM_THREADS")}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
