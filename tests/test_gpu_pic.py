"""GPU parity of PIC/FLIP against the oracle (kern/pic.py, fb/solvers/pic_flip.py).

Bar: seeding (PCG32 stream 71, row-major fluid cells) is bit-exact; G2P uses the reference's
fixed 3 x 3 FP64 order and is bit-exact; P2G sums each face's contributions in cell order
instead of particle order, so FP64 faces agree to 1e-13 relative and FP32 faces to 1e-5,
with identical unsupported-face counts. Whole advances (collisions, P2G, viscosity,
projection, FLIP blend, advection, swept collisions, respawn, population maintenance) agree
with the FP64 oracle to 1e-9 relative; the device PCG32 stream ends in the oracle's state.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import pic as P
from oracle.mac import Dom, cell_to_faces
from oracle.pic import PicFlipOracle
from tests.helpers import make_scenario

pytestmark = pytest.mark.gpu


def _lib():
    from foilbench_amd import _native as N

    return N, N.Context.default()


def _p2f(pos, vel, fu, fv, nx, ny, x0, y0, dx, dy, px, py):
    N, ctx = _lib()
    prec = N.FB_F64 if pos.dtype == np.float64 else N.FB_F32
    u = np.empty_like(fu)
    v = np.empty_like(fv)
    frac = C.c_double()
    N.check(ctx.lib.fb_pic_particle_to_faces(ctx.handle, prec, pos.shape[0], N.ptr(pos), N.ptr(vel), N.ptr(fu),
                                             N.ptr(fv), nx, ny, x0, y0, dx, dy, int(px), int(py), N.ptr(u),
                                             N.ptr(v), C.byref(frac)))
    return u, v, frac.value


def _f2p(u, v, pos, nx, ny, x0, y0, dx, dy, px, py):
    N, ctx = _lib()
    prec = N.FB_F64 if pos.dtype == np.float64 else N.FB_F32
    out = np.empty_like(pos)
    N.check(ctx.lib.fb_pic_faces_to_particle(ctx.handle, prec, pos.shape[0], N.ptr(u), N.ptr(v), N.ptr(pos), nx,
                                             ny, x0, y0, dx, dy, int(px), int(py), N.ptr(out)))
    return out


def _particles(rng, n, nx, ny, x0, y0, dx, dy, dtype, spill=0.0):
    """Uniform particles over the box, `spill` cells beyond each edge."""
    x = x0 - spill * dx + rng.random(n) * (nx + 2 * spill) * dx
    y = y0 - spill * dy + rng.random(n) * (ny + 2 * spill) * dy
    return np.ascontiguousarray(np.stack((x, y), axis=1).astype(dtype))


@pytest.mark.parametrize("periodic", [(False, False), (True, False), (False, True), (True, True)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_transfers_match_oracle(periodic, dtype):
    px, py = periodic
    rng = np.random.default_rng(11)
    nx, ny, x0, y0, dx, dy = 23, 17, -1.25, -0.75, 0.125, 0.1
    # sparse enough that some faces are unsupported; a few particles beyond the edges
    pos = _particles(rng, 700, nx, ny, x0, y0, dx, dy, dtype, spill=1.5)
    vel = rng.normal(size=(700, 2)).astype(dtype)
    fu = rng.normal(size=(ny, nx + 1)).astype(dtype)
    fv = rng.normal(size=(ny + 1, nx)).astype(dtype)
    u, v, frac = _p2f(pos, vel, fu, fv, nx, ny, x0, y0, dx, dy, px, py)
    eu, ev, efrac = P.particle_to_faces(pos, vel, fu, fv, x0, y0, dx, dy, px, py)
    assert frac == efrac
    tol = 1e-13 if dtype == np.float64 else 1e-5
    np.testing.assert_allclose(u, eu, rtol=tol, atol=tol)
    np.testing.assert_allclose(v, ev, rtol=tol, atol=tol)
    if px:
        np.testing.assert_array_equal(u[:, -1], u[:, 0])
    if py:
        np.testing.assert_array_equal(v[-1, :], v[0, :])
    g = _f2p(eu, ev, pos, nx, ny, x0, y0, dx, dy, px, py)
    np.testing.assert_array_equal(g, P.faces_to_particle(eu, ev, pos, x0, y0, dx, dy, px, py))


def test_transfer_edge_cases():
    # empty particle set: every face falls back and counts as unsupported
    nx, ny = 8, 6
    fu = np.arange(ny * (nx + 1), dtype=np.float64).reshape(ny, nx + 1)
    fv = -np.arange((ny + 1) * nx, dtype=np.float64).reshape(ny + 1, nx)
    empty = np.zeros((0, 2))
    u, v, frac = _p2f(empty, empty, fu, fv, nx, ny, 0.0, 0.0, 0.5, 0.5, False, False)
    np.testing.assert_array_equal(u, fu)
    np.testing.assert_array_equal(v, fv)
    assert frac == 1.0
    # all particles in one cell (a clump far above the usual occupancy)
    pos = np.full((300, 2), 1.3) + np.linspace(0.0, 0.2, 300)[:, None]
    vel = np.stack((np.linspace(-1, 1, 300), np.linspace(2, 3, 300)), axis=1)
    u, v, frac = _p2f(pos, vel, fu, fv, nx, ny, 0.0, 0.0, 0.5, 0.5, False, False)
    eu, ev, efrac = P.particle_to_faces(pos, vel, fu, fv, 0.0, 0.0, 0.5, 0.5)
    assert frac == efrac
    np.testing.assert_allclose(u, eu, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(v, ev, rtol=1e-13, atol=1e-13)


def test_constant_velocity_is_preserved_at_particle_positions():
    # tests/unit/test_pic_transfers.py:198-258 on the device
    ny, nx = 6, 8
    x0, y0, dx, dy = -1.0, -0.75, 0.25, 0.25
    pos = np.asarray([[x0 + (x + 0.25) * dx, y0 + (y + 0.75) * dy] for y in range(ny) for x in range(nx)])
    vel = np.broadcast_to(np.asarray([0.75, -0.25]), pos.shape).copy()
    fu = np.full((ny, nx + 1), 0.75)
    fv = np.full((ny + 1, nx), -0.25)
    u, v, _ = _p2f(pos, vel, fu, fv, nx, ny, x0, y0, dx, dy, True, True)
    g = _f2p(u, v, pos, nx, ny, x0, y0, dx, dy, True, True)
    np.testing.assert_allclose(u, fu, rtol=1e-12)
    np.testing.assert_allclose(v, fv, rtol=1e-12)
    np.testing.assert_allclose(g, vel, rtol=1e-12)


def _solver(**kw):
    from foilbench_amd.solvers.pic_flip import PicFlipSolver

    return PicFlipSolver(**kw)


def _foil(sc):
    from foilbench_amd.core.geometry import NacaFoil

    return NacaFoil(sc.foil)


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_seeding_is_bit_exact(precision):
    sc = make_scenario(resolution=(40, 20), bounds=((-1.5, 2.5), (-1.0, 1.0)), controls=((0.0, -10.0),),
                       naca="4412", precision=precision)
    gpu = _solver()
    gpu.initialize(sc, _foil(sc), 42)
    orc = PicFlipOracle()
    orc.initialize(sc, seed=42)
    np.testing.assert_array_equal(gpu.solid_mask(), orc.solid)
    pos, vel = gpu.particles()
    np.testing.assert_array_equal(pos, orc.positions)
    np.testing.assert_array_equal(vel, orc.velocities)
    assert gpu.rng_words() == (orc.rng.state, orc.rng.inc)
    assert gpu._unsupported_face_fraction == orc.unsupported
    np.testing.assert_array_equal(gpu.grid_velocity(), orc.grid)


CASES = {
    # static cambered foil with open boundaries: escapes at the outlet, respawn at the inlet
    "static": dict(resolution=(48, 24), bounds=((-1.5, 2.5), (-1.0, 1.0)), controls=((0.0, -10.0),),
                   naca="4412", output_dt=0.02, reynolds=500.0, pic_population_interval=2),
    # rotating foil: swept collisions, per-substep masks, moving-wall velocity
    "moving": dict(resolution=(48, 24), bounds=((-1.5, 2.5), (-1.0, 1.0)), controls=((0.0, 0.0), (0.2, 25.0)),
                   naca="0012", output_dt=0.02, reynolds=500.0, pic_population_interval=2),
    # periodic channel: wrap instead of respawn, duplicate periodic faces
    "poiseuille": dict(resolution=(32, 16), initial_condition="poiseuille", foil_in_domain=False,
                       periodic_axes=("x",), bounds=((0.0, 2.0), (-0.5, 0.5)), output_dt=0.02,
                       controls=((0.0, 0.0),), reynolds=100.0),
}


def _rel(a, b):
    return float(np.linalg.norm((np.asarray(a, np.float64) - b).ravel()) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.mark.parametrize("case", list(CASES))
def test_advance_matches_oracle_fp64(case):
    sc = make_scenario(precision="float64", **CASES[case])
    gpu = _solver()
    gpu.initialize(sc, _foil(sc), 7)
    orc = PicFlipOracle()
    orc.initialize(sc, seed=7)
    t = 0.0
    for _ in range(4):
        c = sc.control_at(t + sc.output_dt)
        rg = gpu.advance(c, sc.output_dt)
        ro = orc.advance(c, sc.output_dt)
        t += sc.output_dt
        assert rg.substeps == ro["substeps"]
        assert abs(rg.evidence["pressure_iterations"] - ro["pressure_iterations"]) <= 1
        assert abs(rg.evidence["viscosity_iterations"] - ro["viscosity_iterations"]) <= 1
        assert rg.evidence["unsupported_face_fraction"] == pytest.approx(ro["unsupported_face_fraction"], abs=1e-12)
    pos, vel = gpu.particles()
    assert pos.shape == orc.positions.shape
    assert _rel(pos, orc.positions) < 1e-9
    assert _rel(vel, orc.velocities) < 1e-9
    assert _rel(gpu.grid_velocity(), orc.grid) < 1e-9
    assert gpu.rng_words() == (orc.rng.state, orc.rng.inc)


def test_advance_fp32_tracks_oracle():
    sc = make_scenario(precision="float32", **CASES["static"])
    gpu = _solver()
    gpu.initialize(sc, _foil(sc), 3)
    orc = PicFlipOracle()
    orc.initialize(sc, seed=3)
    t = 0.0
    for _ in range(2):
        c = sc.control_at(t + sc.output_dt)
        gpu.advance(c, sc.output_dt)
        orc.advance(c, sc.output_dt)
        t += sc.output_dt
    pos, vel = gpu.particles()
    assert pos.dtype == np.float32
    assert _rel(pos, orc.positions) < 1e-5
    assert _rel(gpu.grid_velocity(), orc.grid) < 1e-4


def test_rollback_restores_particles_and_stream_bitwise():
    from foilbench_amd.core.models import NumericalFailure

    sc = make_scenario(precision="float64", **dict(CASES["static"], mac_maximum_divergence_linf=0.0))
    gpu = _solver()
    gpu.initialize(sc, _foil(sc), 1)
    pos0, vel0 = gpu.particles()
    grid0 = gpu.grid_velocity()
    words0 = gpu.rng_words()
    with pytest.raises(NumericalFailure) as err:
        gpu.advance(sc.control_at(sc.output_dt), sc.output_dt)
    assert err.value.reason == "postcondition_failure"
    pos1, vel1 = gpu.particles()
    np.testing.assert_array_equal(pos1, pos0)
    np.testing.assert_array_equal(vel1, vel0)
    np.testing.assert_array_equal(gpu.grid_velocity(), grid0)
    assert gpu.rng_words() == words0
    assert gpu.state_revision == 0


def test_particles_stay_outside_the_foil_and_counts_are_reported():
    sc = make_scenario(precision="float64", **CASES["moving"])
    gpu = _solver()
    gpu.initialize(sc, _foil(sc), 5)
    t = 0.0
    for _ in range(3):
        c = sc.control_at(t + sc.output_dt)
        rep = gpu.advance(c, sc.output_dt)
        t += sc.output_dt
        assert rep.evidence["unresolved_solid_particles"] == 0
        assert rep.evidence["particle_count"] == gpu.particles()[0].shape[0]
    pos, _ = gpu.particles()
    inside = _foil(sc).contains(pos, gpu._control.angle_degrees)
    assert not np.any(inside)
    d = gpu.diagnostics()
    assert d.values["particles_inside_solid"] == 0.0
