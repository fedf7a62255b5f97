"""Short workload for rocprofv3 --pmc passes: a 256 MiB device copy (counter calibration)
followed by lattice steps of the LBM step kernel with its default launch shape.
``python -m foilbench_amd.benchmark.pmc_lbm [steps] [nx] [ny] [form] [rows nt]`` (defaults: 10 steps of the
1024 x 512 FP32 lattice; form is one of f64, f32, f32-mixed, f16-mixed; `rows nt` fix the interior rows per
thread and the nontemporal population traffic of the single-step kernel, otherwise the lattice's defaults)."""

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
    if len(sys.argv) > 6:
        prec = _PREC[form]
        N.check(ctx.lib.fb_tune_set(N.FB_TUNE_LBM_FUSE, nx, ny, prec, 1, 0))
        N.check(ctx.lib.fb_tune_set(N.FB_TUNE_LBM_ROWS, nx, ny, prec, 1, int(sys.argv[5])))
        N.check(ctx.lib.fb_tune_set(N.FB_TUNE_LBM_NT, nx, ny, prec, 1, int(sys.argv[6])))
    ms = C.c_double()
    N.check(ctx.lib.fb_copy_benchmark(ctx.handle, 256 * 1024 * 1024, 2, C.byref(ms)))
    lat = Lattice(LatticeCase(nx=nx, ny=ny), _PREC[form], ctx)
    lat.run(steps)
    ctx.synchronize()
    print(f"copy 256 MiB x3 launches, lbm {form} {nx}x{ny} {steps} steps done", flush=True)


if __name__ == "__main__":
    main()
