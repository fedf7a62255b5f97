"""ctypes binding of oracle/lbm_oracle.c: ``kernel_steps`` of oracle/lbm.py as C loops (test infrastructure).

Same arguments and results as oracle.lbm.kernel_steps, bit for bit (tests/test_oracle_lbm_c.py), at C speed
and on several threads; used as the checker of the GPU lattice at full sizes and horizons and as the CPU
baseline of the lattice step.
"""

from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liblbm_oracle.so"


class _Params(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("periodic_x", C.c_int), ("periodic_y", C.c_int),
                ("channel", C.c_int), ("omega_plus", C.c_double), ("omega_minus", C.c_double),
                ("u_lattice", C.c_double), ("target_ux", C.c_double), ("target_uy", C.c_double),
                ("feq_b", C.c_double * 9), ("x0", C.c_double), ("y0", C.c_double), ("dx", C.c_double),
                ("dy", C.c_double), ("pivot_x", C.c_double), ("pivot_y", C.c_double), ("omega_rad", C.c_double),
                ("scale", C.c_double)]


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m oracle.build`")
        lib = C.CDLL(str(LIB_PATH))
        lib.lbm_oracle_steps.restype = C.c_int
        lib.lbm_oracle_steps.argtypes = [C.POINTER(_Params), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_int, C.c_int]
        _lib = lib
    return _lib


def kernel_steps(f, solid, sdf, omega_plus, omega_minus, u_lattice, target, feq_b, sponge_s, outlet,
                 bounds, dx, dy, pivot, omega_deg, scale, periodic_x, periodic_y, channel, steps, threads=0):
    """oracle.lbm.kernel_steps in C: returns (f, outlet) after `steps` lattice steps (new arrays)."""
    lib = _load()
    f = np.ascontiguousarray(f).copy()
    ny, nx, _ = f.shape
    out = np.ascontiguousarray(np.asarray(outlet, dtype=f.dtype).reshape(ny, 9)).copy()
    solid = np.ascontiguousarray(solid, dtype=np.uint8)
    sdf = np.ascontiguousarray(sdf, dtype=np.float64)
    sp = np.ascontiguousarray(np.broadcast_to(sponge_s, (ny, nx)), dtype=np.float64)
    p = _Params()
    p.nx, p.ny = nx, ny
    p.periodic_x, p.periodic_y, p.channel = int(bool(periodic_x)), int(bool(periodic_y)), int(bool(channel))
    p.omega_plus, p.omega_minus, p.u_lattice = float(omega_plus), float(omega_minus), float(u_lattice)
    p.target_ux, p.target_uy = float(target[0]), float(target[1])
    fb = np.asarray(feq_b, dtype=f.dtype).reshape(9)
    for d in range(9):
        p.feq_b[d] = float(fb[d])
    p.x0, p.y0, p.dx, p.dy = float(bounds[0][0]), float(bounds[1][0]), float(dx), float(dy)
    p.pivot_x, p.pivot_y = float(pivot[0]), float(pivot[1])
    p.omega_rad = math.radians(omega_deg) if omega_deg else 0.0
    p.scale = float(scale)
    storage = {np.dtype(np.float64): 8, np.dtype(np.float32): 4}[f.dtype]
    vp = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    rc = lib.lbm_oracle_steps(C.byref(p), storage, vp(f), vp(out), vp(solid), vp(sdf), vp(sp), int(steps),
                              int(threads))
    if rc:
        raise RuntimeError(f"lbm_oracle_steps failed ({rc})")
    return f, out.reshape(ny, 1, 9)
