"""Pin the PIC/FLIP oracle (oracle/pic.py): per-particle loops restating the numba kernels
on tiny inputs, and the reference's transfer and population tests."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import pic as P
from oracle.mac import Dom
from oracle.rng import PCG32


def _w(d):
    a = abs(d)
    if a < 0.5:
        return 0.75 - a * a
    if a < 1.5:
        return 0.5 * (1.5 - a) ** 2
    return 0.0


def _loop_scatter(pos, vals, x0, y0, dx, dy, W, H, gxo, gyo, px, py, dupx, dupy, fb):
    w = np.zeros((H, W), pos.dtype)
    m = np.zeros((H, W), pos.dtype)
    uw = W - 1 if dupx else W
    uh = H - 1 if dupy else H
    for p in range(pos.shape[0]):
        gx = (float(pos[p, 0]) - x0) / dx + gxo
        gy = (float(pos[p, 1]) - y0) / dy + gyo
        bx = int(np.floor(gx - 0.5))
        by = int(np.floor(gy - 0.5))
        for oy in range(3):
            sy = by + oy
            ty = sy % uh if py else sy
            if ty < 0 or ty >= H:
                continue
            wy = _w(gy - sy)
            for ox in range(3):
                sx = bx + ox
                tx = sx % uw if px else sx
                if tx < 0 or tx >= W:
                    continue
                wt = wy * _w(gx - sx)
                # numba: f32 element + f64 value computed in FP64, rounded on store
                w[ty, tx] = float(w[ty, tx]) + wt
                m[ty, tx] = float(m[ty, tx]) + wt * float(vals[p])
    out = np.where(w > 1e-12, m / np.where(w > 1e-12, w, 1), fb).astype(pos.dtype)
    if dupx:
        out[:, W - 1] = out[:, 0]
    if dupy:
        out[H - 1, :] = out[0, :]
    return out


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_transfers_match_per_particle_loops(periodic, dtype):
    rng = np.random.default_rng(5)
    nx, ny = 7, 5
    x0, y0, dx, dy = -1.0, -0.5, 0.25, 0.2
    pos = np.stack((x0 + rng.random(30) * nx * dx, y0 + rng.random(30) * ny * dy), axis=1).astype(dtype)
    vel = rng.normal(size=(30, 2)).astype(dtype)
    fu = rng.normal(size=(ny, nx + 1)).astype(dtype)
    fv = rng.normal(size=(ny + 1, nx)).astype(dtype)
    u, v, uns = P.particle_to_faces(pos, vel, fu, fv, x0, y0, dx, dy, periodic, periodic)
    eu = _loop_scatter(pos, vel[:, 0], x0, y0, dx, dy, nx + 1, ny, 0.0, -0.5, periodic, periodic, periodic, False, fu)
    ev = _loop_scatter(pos, vel[:, 1], x0, y0, dx, dy, nx, ny + 1, -0.5, 0.0, periodic, periodic, False, periodic, fv)
    np.testing.assert_array_equal(u, eu)
    np.testing.assert_array_equal(v, ev)
    # gather: per-particle fixed 3x3 loop
    g = P.faces_to_particle(u, v, pos, x0, y0, dx, dy, periodic, periodic)
    for p in range(pos.shape[0]):
        for comp, (f, gxo, gyo, dupx, dupy) in enumerate(((u, 0.0, -0.5, periodic, False),
                                                          (v, -0.5, 0.0, False, periodic))):
            H, W = f.shape
            uw = W - 1 if dupx else W
            uh = H - 1 if dupy else H
            gx = (float(pos[p, 0]) - x0) / dx + gxo
            gy = (float(pos[p, 1]) - y0) / dy + gyo
            bx, by = int(np.floor(gx - 0.5)), int(np.floor(gy - 0.5))
            val = ws = 0.0
            for oy in range(3):
                sy = by + oy
                ty = sy % uh if periodic else sy
                if ty < 0 or ty >= H:
                    continue
                for ox in range(3):
                    sx = bx + ox
                    tx = sx % uw if periodic else sx
                    if tx < 0 or tx >= W:
                        continue
                    wt = _w(gy - sy) * _w(gx - sx)
                    val += wt * float(f[ty, tx])
                    ws += wt
            assert g[p, comp] == dtype(val / max(ws, 1e-12))


def test_quadratic_bspline_partition_of_unity():
    for frac in np.linspace(0.0, 0.999, 37):
        g = 10.0 + frac
        b = int(np.floor(g - 0.5))
        assert sum(_w(g - (b + k)) for k in range(3)) == pytest.approx(1.0, abs=1e-15)
        np.testing.assert_allclose(P.bspline(np.asarray([g - b, g - b - 1, g - b - 2])).sum(), 1.0, atol=1e-15)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_direct_mac_transfer_preserves_constant_velocity(dtype):
    # tests/unit/test_pic_transfers.py:198-258
    ny, nx = 6, 8
    x0, y0, dx, dy = -1.0, -0.75, 0.25, 0.25
    pos = np.asarray([[x0 + (x + 0.25) * dx, y0 + (y + 0.75) * dy] for y in range(ny) for x in range(nx)], dtype)
    vel = np.broadcast_to(np.asarray([0.75, -0.25], dtype), pos.shape).copy()
    fu = np.full((ny, nx + 1), 0.75, dtype)
    fv = np.full((ny + 1, nx), -0.25, dtype)
    u, v, uns = P.particle_to_faces(pos, vel, fu, fv, x0, y0, dx, dy, True, True)
    g = P.faces_to_particle(u, v, pos, x0, y0, dx, dy, True, True)
    np.testing.assert_allclose(u, fu, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(v, fv, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(g, vel, rtol=1e-6, atol=1e-6)
    np.testing.assert_array_equal(u[:, 0], u[:, -1])
    np.testing.assert_array_equal(v[0, :], v[-1, :])
    assert 0.0 <= uns < 1.0


def test_population_maintenance_repairs_holes_and_clumps():
    # tests/unit/test_particle_population.py:20-40
    dom = Dom(((0.0, 1.0), (0.0, 1.0)), 4, 4)
    solid = np.zeros((4, 4), bool)
    solid[1, 2] = True
    n = 4 * int(np.count_nonzero(~solid))
    a = np.full((n, 2), 0.125, np.float32)
    b = a.copy()
    ma, _ = P.maintain_particle_population(a, solid, dom, PCG32(17))
    mb, _ = P.maintain_particle_population(b, solid, dom, PCG32(17))
    np.testing.assert_array_equal(ma, mb)
    np.testing.assert_array_equal(a, b)
    c = P.particle_cell_counts(a, dom).reshape(4, 4)
    assert np.all(c[~solid] == 4) and np.all(c[solid] == 0)


def test_population_maintenance_inactive_within_limits():
    dom = Dom(((0.0, 1.0), (0.0, 1.0)), 4, 4)
    solid = np.zeros((4, 4), bool)
    pos = np.asarray([[(cx + o) * 0.25, (cy + o) * 0.25] for cy in range(4) for cx in range(4)
                      for o in (0.35, 0.45, 0.55, 0.65)], np.float32)
    before = pos.copy()
    moved, _ = P.maintain_particle_population(pos, solid, dom, PCG32(3))
    assert moved.size == 0
    np.testing.assert_array_equal(pos, before)


def test_seeding_uses_pcg32_stream_71_in_row_major_order(golden):
    dom = Dom(((0.0, 1.0), (0.0, 0.5)), 4, 2)
    solid = np.zeros((2, 4), bool)
    solid[0, 1] = True
    pos = P.seed_particles(solid, dom, PCG32(42, stream=71), np.float64, per_cell=4)
    assert pos.shape == (28, 2)
    draws = PCG32(42, stream=71).random((56,))
    # particle 4 is the first one of the second fluid cell (x=2, y=0)
    # the jitter is cast to the scenario dtype (FP64 here) before 0.8 * jitter; int64 + 0.1 is FP64
    jit = draws.astype(np.float64)
    ex = 0.0 + (np.int64(2) + 0.1 + 0.8 * jit[8]) * 0.25
    ey = 0.0 + (np.int64(0) + 0.1 + 0.8 * jit[9]) * 0.25
    np.testing.assert_array_equal(pos[4], [ex, ey])
    cells = P.particle_cell_ids(pos, dom)
    assert np.all(np.bincount(cells, minlength=8)[~solid.ravel()] == 4)
