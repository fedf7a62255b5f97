"""Short workload for rocprofv3 --pmc passes: a 256 MiB device copy (counter calibration)
followed by lattice steps of the LBM step kernel with its default launch shape.
``python -m foilbench_amd.benchmark.pmc_lbm [steps] [nx] [ny] [form]`` (defaults: 10 steps of the
1024 x 512 FP32 lattice; form is one of f64, f32, f32-mixed, f16-mixed)."""

from __future__ import annotations

import ctypes as C
import sys

from .. import _native as N
from ..tuning import _PREC
from .lbm_harness import Lattice, LatticeCase


def main() -> None:
    steps = int(sys.argv[1]) if len(sys.argv) > 1 else 10
    nx = int(sys.argv[2]) if len(sys.argv) > 2 else 1024
    ny = int(sys.argv[3]) if len(sys.argv) > 3 else 512
    form = sys.argv[4] if len(sys.argv) > 4 else "f32"
    ctx = N.Context.default()
    ms = C.c_double()
    N.check(ctx.lib.fb_copy_benchmark(ctx.handle, 256 * 1024 * 1024, 2, C.byref(ms)))
    lat = Lattice(LatticeCase(nx=nx, ny=ny), _PREC[form], ctx)
    lat.run(steps)
    ctx.synchronize()
    print(f"copy 256 MiB x3 launches, lbm {form} {nx}x{ny} {steps} steps done", flush=True)


if __name__ == "__main__":
    main()
