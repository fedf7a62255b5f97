"""Builds the oracle's C restatements (test infrastructure): oracle/lbm_oracle.c -> oracle/liblbm_oracle.so.

``python -m oracle.build`` (also run by ``__graft_entry__.build()``). gcc with -O2 -ffp-contract=off, so the
C loops keep the NumPy restatement's rounding, and OpenMP over rows.
"""

from __future__ import annotations

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB = HERE / "liblbm_oracle.so"


def build(verbose: bool = True) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC", "-shared", "-std=c11",
           str(HERE / "lbm_oracle.c"), "-o", str(tmp), "-lm"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"gcc failed for lbm_oracle.c:\n{res.stderr[-4000:]}")
    tmp.replace(LIB)
    if verbose:
        print(f"built {LIB}", flush=True)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
