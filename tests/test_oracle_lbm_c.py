"""The C restatement of the lattice steps (oracle/lbm_oracle.c via oracle/lbm_c.py) against the NumPy one
(oracle/lbm.py kernel_steps): identical populations and outlet history, bit for bit, in FP64 and FP32 storage,
on a static foil with open boundaries, a rotating wall, a periodic box and a channel, for 1 and 4 threads."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import lbm as L
from oracle import lbm_c
from oracle.geometry import NacaFoil, cell_centers

CASES = {
    "open": dict(nx=48, ny=24, periodic=(False, False), channel=False, omega=0.0, naca="2412", angle=-12.0),
    "rotating": dict(nx=40, ny=24, periodic=(False, False), channel=False, omega=40.0, naca="4412", angle=-6.0),
    "periodic": dict(nx=32, ny=20, periodic=(True, True), channel=False, omega=0.0, naca="0012", angle=0.0),
    "channel": dict(nx=36, ny=18, periodic=(True, False), channel=True, omega=0.0, naca="0012", angle=5.0),
}


def _setup(case, dtype):
    c = CASES[case]
    nx, ny = c["nx"], c["ny"]
    bounds = ((-1.6, 4.8), (-1.6, 1.6))
    dx, dy = 6.4 / nx, 3.2 / ny
    foil = NacaFoil(c["naca"], 1.0, (0.0, 0.0))
    sdf = foil.signed_distance(cell_centers(bounds, nx, ny).reshape(-1, 2), c["angle"]).reshape(ny, nx)
    solid = sdf <= 0.0
    rng = np.random.default_rng(3)
    u = 0.05
    vel = np.zeros((ny, nx, 2), dtype)
    vel[..., 0] = u
    vel += (0.01 * rng.standard_normal(vel.shape)).astype(dtype)
    rho = (1.0 + 0.01 * rng.standard_normal((ny, nx))).astype(dtype)
    f = L.equilibrium(rho, vel)
    target = tuple(np.asarray([u, 0.0], dtype=dtype))
    feqb = L.equilibrium(np.ones((1, 1), dtype), np.asarray(target, dtype).reshape(1, 1, 2))[0, 0]
    sp = L.sponge_strength_profile(nx, ny, c["periodic"][0], c["periodic"][1], c["channel"])
    outlet = f[:, -1:].copy()
    nu = 0.01
    tp = 0.5 + 3 * nu
    args = (solid, sdf, 1.0 / tp, 1.0 / (0.5 + (3 / 16) / (tp - 0.5)), u, target, feqb, sp, outlet, bounds, dx, dy,
            (0.0, 0.0), c["omega"], u, c["periodic"][0], c["periodic"][1], c["channel"])
    return f, args


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("case", list(CASES))
def test_c_lattice_steps_equal_numpy(case, dtype):
    f, args = _setup(case, dtype)
    steps = 7
    fn, on = L.kernel_steps(f.copy(), *args, steps)
    for threads in (1, 4):
        fc, oc = lbm_c.kernel_steps(f, *args, steps, threads=threads)
        assert fc.dtype == fn.dtype
        np.testing.assert_array_equal(fc, fn)
        np.testing.assert_array_equal(oc.reshape(-1, 9), on.reshape(-1, 9))
    assert np.isfinite(fn).all()
