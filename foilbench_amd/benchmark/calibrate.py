"""HBM calibration: a plain device copy of known size (for rocprof byte counters and the
achievable bandwidth). ``python -m foilbench_amd.benchmark.calibrate [MiB] [reps]``"""

from __future__ import annotations

import ctypes as C
import json
import sys

from .. import _native as N


def main() -> None:
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    reps = int(sys.argv[2]) if len(sys.argv) > 2 else 20
    ctx = N.Context.default()
    ms = C.c_double()
    nbytes = mib * 1024 * 1024
    N.check(ctx.lib.fb_copy_benchmark(ctx.handle, nbytes, reps, C.byref(ms)))
    gbs = 2 * nbytes / (ms.value * 1e-3) / 1e9
    print(json.dumps({"copy_bytes": nbytes, "read_plus_write_bytes": 2 * nbytes, "ms_per_copy": ms.value,
                      "GBps": gbs}), flush=True)


if __name__ == "__main__":
    main()
