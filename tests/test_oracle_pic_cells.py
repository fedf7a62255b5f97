"""Pin the cell-centred transfers of the oracle (oracle/pic.py particle_to_grid / grid_to_particle, restating
kern/pic.py:22-102): the reference's own unit cases (tests/unit/test_pic_transfers.py:22-196), closed forms of
the quadratic B-spline (one particle's 3 x 3 support and weights, clamping at the edges, exact reproduction of
linear fields by the gather, conservation of weight and momentum by the scatter), and a per-particle loop of the
numba kernels' arithmetic for the accumulation order and rounding."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import pic as P


def _w(d):
    a = abs(d)
    if a < 0.5:
        return 0.75 - a * a
    if a < 1.5:
        return 0.5 * (1.5 - a) ** 2
    return 0.0


def _reference_grid_to_particle(grid, positions, x0, y0, dx, dy):
    """The reference test's own NumPy restatement (tests/unit/test_pic_transfers.py:22-51): clamped targets,
    division by the weight sum."""
    ny, nx, _ = grid.shape
    out = np.empty_like(positions)
    for p, pos in enumerate(positions):
        gx = (pos[0] - x0) / dx - 0.5
        gy = (pos[1] - y0) / dy - 0.5
        bx, by = int(np.floor(gx - 0.5)), int(np.floor(gy - 0.5))
        val = np.zeros(2)
        ws = 0.0
        for oy in range(3):
            sy = by + oy
            ty = min(max(sy, 0), ny - 1)
            for ox in range(3):
                sx = bx + ox
                tx = min(max(sx, 0), nx - 1)
                w = _w(gy - sy) * _w(gx - sx)
                val += w * grid[ty, tx]
                ws += w
        out[p] = val / ws
    return out


@pytest.mark.parametrize(("shape", "dtype"), [((4, 5), np.float32), ((6, 8), np.float64)])
def test_grid_to_particle_matches_the_reference_test(shape, dtype):
    # tests/unit/test_pic_transfers.py:54-78
    ny, nx = shape
    x0, y0, dx, dy = -1.25, -0.75, 0.2, 0.15
    grid = np.arange(ny * nx * 2, dtype=dtype).reshape(ny, nx, 2) / 17.0
    pos = np.asarray([[x0 + 0.01 * dx, y0 + 0.02 * dy], [x0 + 1.3 * dx, y0 + 2.7 * dy],
                      [x0 + (nx - 0.01) * dx, y0 + (ny - 0.02) * dy]], dtype=dtype)
    got = P.grid_to_particle(grid, pos, x0, y0, dx, dy)
    assert got.dtype == dtype
    np.testing.assert_allclose(got, _reference_grid_to_particle(grid, pos, x0, y0, dx, dy), rtol=2e-6, atol=2e-6)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_grid_to_particle_preserves_constant_velocity(dtype):
    # tests/unit/test_pic_transfers.py:81-98
    grid = np.empty((5, 7, 2), dtype=dtype)
    grid[...] = np.asarray([1.25, -0.375], dtype=dtype)
    pos = np.asarray([[-0.99, -0.49], [-0.25, 0.0], [0.99, 0.49]], dtype=dtype)
    got = P.grid_to_particle(grid, pos, -1.0, -0.5, 2.0 / 7.0, 1.0 / 5.0)
    np.testing.assert_allclose(got, np.broadcast_to(np.asarray([1.25, -0.375], dtype), got.shape), rtol=1e-6,
                               atol=1e-6)


def test_periodic_transfer_wraps_quadratic_support():
    # tests/unit/test_pic_transfers.py:143-196
    grid = np.arange(6 * 8 * 2, dtype=np.float32).reshape(6, 8, 2)
    x0, y0, dx, dy = -1.0, -0.75, 0.25, 0.25
    width, height = grid.shape[1] * dx, grid.shape[0] * dy
    pos = np.asarray([[x0 - 0.07, y0 + 0.31], [x0 + width - 0.07, y0 + 0.31], [x0 + 0.43, y0 - 0.09],
                      [x0 + 0.43, y0 + height - 0.09]], dtype=np.float32)
    g = P.grid_to_particle(grid, pos, x0, y0, dx, dy, True, True)
    np.testing.assert_allclose(g[0], g[1], rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(g[2], g[3], rtol=1e-6, atol=1e-6)
    vel = np.asarray([[0.75, -0.25], [0.75, -0.25]], dtype=np.float32)
    left = P.particle_to_grid(pos[:1], vel[:1], x0, y0, dx, dy, 8, 6, (0.0, 0.0), True, True)
    right = P.particle_to_grid(pos[1:2], vel[1:2], x0, y0, dx, dy, 8, 6, (0.0, 0.0), True, True)
    np.testing.assert_allclose(left, right, rtol=1e-6, atol=1e-6)


def test_one_particle_at_a_cell_centre_fills_its_3x3_support():
    # at a cell centre the quadratic B-spline weights are 1/8, 3/4, 1/8 per axis; two particles in the same
    # support give the weight-averaged velocity; every other cell takes the freestream
    nx, ny, dx, dy = 9, 7, 0.5, 0.25
    pos = np.asarray([[(4 + 0.5) * dx, (3 + 0.5) * dy]])
    grid = P.particle_to_grid(pos, np.asarray([[2.0, -1.0]]), 0.0, 0.0, dx, dy, nx, ny, (7.0, 8.0))
    support = np.zeros((ny, nx), bool)
    support[2:5, 3:6] = True
    np.testing.assert_array_equal(grid[support], np.broadcast_to([2.0, -1.0], (9, 2)))
    np.testing.assert_array_equal(grid[~support], np.broadcast_to([7.0, 8.0], ((~support).sum(), 2)))
    pos2 = np.asarray([[(4 + 0.5) * dx, (3 + 0.5) * dy], [(5 + 0.5) * dx, (3 + 0.5) * dy]])
    grid = P.particle_to_grid(pos2, np.asarray([[2.0, 0.0], [6.0, 0.0]]), 0.0, 0.0, dx, dy, nx, ny, (0.0, 0.0))
    # cell (3, 5): weights 3/4 * 1/8 from particle 1 and 3/4 * 3/4 from particle 2
    w1, w2 = 0.75 * 0.125, 0.75 * 0.75
    assert grid[3, 5, 0] == pytest.approx((w1 * 2.0 + w2 * 6.0) / (w1 + w2), rel=1e-15)
    # cell (3, 6): only particle 2 (weight 3/4 * 1/8)
    assert grid[3, 6, 0] == 6.0


def test_clamping_piles_outside_weights_onto_the_edge_cells():
    nx, ny, dx, dy = 6, 5, 1.0, 1.0
    # 2.9 cells left of the box, in row 2: every x-target clamps to column 0
    grid = P.particle_to_grid(np.asarray([[-2.9, 2.5]]), np.asarray([[1.5, 0.5]]), 0.0, 0.0, dx, dy, nx, ny,
                              (0.0, 0.0))
    touched = np.argwhere(grid[..., 0] != 0.0)
    assert set(map(tuple, touched)) == {(1, 0), (2, 0), (3, 0)}
    # periodic in x instead: the support wraps to columns 2..4 (stencil -4..-2 modulo 6)
    grid = P.particle_to_grid(np.asarray([[-2.9, 2.5]]), np.asarray([[1.5, 0.5]]), 0.0, 0.0, dx, dy, nx, ny,
                              (0.0, 0.0), True, False)
    touched = np.argwhere(grid[..., 0] != 0.0)
    assert {c for _, c in touched} == {2, 3, 4}
    # the gather clamps the same way: a particle far right of the box sees column nx - 1 only
    g = np.zeros((ny, nx, 2))
    g[:, nx - 1] = 4.0
    assert P.grid_to_particle(g, np.asarray([[40.0, 2.5]]), 0.0, 0.0, dx, dy)[0, 0] == 4.0


def test_gather_reproduces_linear_fields_exactly_inside():
    # the quadratic B-spline has zero first moment: sum_i w_i (x_i - x) = 0, so the gather of a linear field
    # at any point at least 1.5 cells inside the box is the field itself
    rng = np.random.default_rng(3)
    nx, ny, x0, y0, dx, dy = 13, 11, -0.7, 0.3, 0.2, 0.1
    xc = x0 + (np.arange(nx) + 0.5) * dx
    yc = y0 + (np.arange(ny) + 0.5) * dy
    X, Y = np.meshgrid(xc, yc)
    grid = np.stack((1.0 + 2.0 * X - 3.0 * Y, -0.5 + 0.25 * X + 4.0 * Y), axis=-1)
    pos = np.stack((x0 + dx * (1.5 + rng.random(200) * (nx - 3)), y0 + dy * (1.5 + rng.random(200) * (ny - 3))),
                   axis=1)
    got = P.grid_to_particle(grid, pos, x0, y0, dx, dy)
    want = np.stack((1.0 + 2.0 * pos[:, 0] - 3.0 * pos[:, 1], -0.5 + 0.25 * pos[:, 0] + 4.0 * pos[:, 1]), axis=1)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-13)


def test_scatter_conserves_weight_and_momentum():
    # each particle's nine weights sum to one, clamped or wrapped: with unit velocities every supported cell is
    # exactly 1, and the weights recovered from a second scatter conserve the particle count and momentum
    rng = np.random.default_rng(4)
    nx, ny, dx, dy = 10, 8, 0.5, 0.5
    pos = np.stack((-1.0 + rng.random(300) * (nx * dx + 2.0), -1.0 + rng.random(300) * (ny * dy + 2.0)), axis=1)
    for px, py in ((False, False), (True, True)):
        tgt, wt = P._cell_stencil(pos, 0.0, 0.0, dx, dy, nx, ny, px, py)
        np.testing.assert_allclose(wt.sum(axis=1), 1.0, atol=1e-15)
        one = P.particle_to_grid(pos, np.ones((300, 2)), 0.0, 0.0, dx, dy, nx, ny, (0.0, 0.0), px, py)
        w = np.bincount(tgt.ravel(), wt.ravel(), nx * ny).reshape(ny, nx)
        np.testing.assert_allclose(one[w > 1e-12], 1.0, atol=1e-14)
        vel = rng.normal(size=(300, 2))
        grid = P.particle_to_grid(pos, vel, 0.0, 0.0, dx, dy, nx, ny, (0.0, 0.0), px, py)
        assert w.sum() == pytest.approx(300.0, rel=1e-13)
        np.testing.assert_allclose((grid * w[..., None]).sum(axis=(0, 1)), vel.sum(axis=0), rtol=1e-12, atol=1e-12)


def _loop_p2g(pos, vel, x0, y0, dx, dy, nx, ny, fs, px, py):
    """particle_to_grid_kernel's loop with its numba arithmetic: f32 element + f64 value in FP64, rounded."""
    dt = pos.dtype
    w = np.zeros((ny, nx), dt)
    m = np.zeros((ny, nx, 2), dt)
    for p in range(pos.shape[0]):
        gx = (float(pos[p, 0]) - x0) / dx - 0.5
        gy = (float(pos[p, 1]) - y0) / dy - 0.5
        bx, by = int(np.floor(gx - 0.5)), int(np.floor(gy - 0.5))
        for oy in range(3):
            sy = by + oy
            ty = sy % ny if py else min(max(sy, 0), ny - 1)
            wy = _w(gy - sy)
            for ox in range(3):
                sx = bx + ox
                tx = sx % nx if px else min(max(sx, 0), nx - 1)
                wt = wy * _w(gx - sx)
                w[ty, tx] = float(w[ty, tx]) + wt
                m[ty, tx, 0] = float(m[ty, tx, 0]) + wt * float(vel[p, 0])
                m[ty, tx, 1] = float(m[ty, tx, 1]) + wt * float(vel[p, 1])
    out = np.empty((ny, nx, 2), dt)
    for y in range(ny):
        for x in range(nx):
            if w[y, x] > 1e-12:
                out[y, x] = m[y, x] / w[y, x]  # dtype quotient
            else:
                out[y, x] = fs
    return out


@pytest.mark.parametrize("periodic", [(False, False), (True, False), (False, True), (True, True)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_cell_transfers_match_per_particle_loops(periodic, dtype):
    rng = np.random.default_rng(6)
    nx, ny = 7, 5
    x0, y0, dx, dy = -1.0, -0.5, 0.25, 0.2
    n = 60
    # particles up to two cells beyond every edge, so clamping and wrapping are exercised
    pos = np.stack((x0 + (rng.random(n) * (nx + 4) - 2) * dx, y0 + (rng.random(n) * (ny + 4) - 2) * dy),
                   axis=1).astype(dtype)
    vel = rng.normal(size=(n, 2)).astype(dtype)
    px, py = periodic
    got = P.particle_to_grid(pos, vel, x0, y0, dx, dy, nx, ny, (0.3, -0.1), px, py)
    np.testing.assert_array_equal(got, _loop_p2g(pos, vel, x0, y0, dx, dy, nx, ny, np.asarray([0.3, -0.1], dtype),
                                                 px, py))
    g = P.grid_to_particle(got, pos, x0, y0, dx, dy, px, py)
    for p in range(n):
        gx = (float(pos[p, 0]) - x0) / dx - 0.5
        gy = (float(pos[p, 1]) - y0) / dy - 0.5
        bx, by = int(np.floor(gx - 0.5)), int(np.floor(gy - 0.5))
        vx = vy = ws = 0.0
        for oy in range(3):
            sy = by + oy
            ty = sy % ny if py else min(max(sy, 0), ny - 1)
            for ox in range(3):
                sx = bx + ox
                tx = sx % nx if px else min(max(sx, 0), nx - 1)
                wt = _w(gy - sy) * _w(gx - sx)
                vx += wt * float(got[ty, tx, 0])
                vy += wt * float(got[ty, tx, 1])
                ws += wt
        inv = 1.0 / max(ws, 1e-12)
        assert g[p, 0] == dtype(vx * inv) and g[p, 1] == dtype(vy * inv)


def test_empty_particle_set_gives_the_freestream():
    grid = P.particle_to_grid(np.zeros((0, 2), np.float32), np.zeros((0, 2), np.float32), 0.0, 0.0, 0.1, 0.1, 4, 3,
                              (1.0, 0.25))
    assert grid.dtype == np.float32
    np.testing.assert_array_equal(grid, np.broadcast_to(np.float32([1.0, 0.25]), (3, 4, 2)))
