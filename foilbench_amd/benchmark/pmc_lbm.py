"""Short workload for rocprofv3 --pmc passes: a 256 MiB device copy (counter calibration)
followed by 10 lattice steps of the 1024x512 FP32 LBM step kernel.
``python -m foilbench_amd.benchmark.pmc_lbm [steps]``"""

from __future__ import annotations

import ctypes as C
import sys

from .. import _native as N
from .lbm_harness import Lattice, LatticeCase


def main() -> None:
    steps = int(sys.argv[1]) if len(sys.argv) > 1 else 10
    ctx = N.Context.default()
    ms = C.c_double()
    N.check(ctx.lib.fb_copy_benchmark(ctx.handle, 256 * 1024 * 1024, 2, C.byref(ms)))
    lat = Lattice(LatticeCase(), N.FB_F32, ctx)
    lat.run(steps)
    ctx.synchronize()
    print(f"copy 256 MiB x3 launches, lbm f32 {steps} steps done", flush=True)


if __name__ == "__main__":
    main()
