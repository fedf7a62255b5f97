"""Pin the LBM oracle with closed forms and the reference's own unit-test restatements
(tests/unit/test_lbm_collision.py, tests/unit/test_lbm_moments.py)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import lbm as L

C = L.C.astype(np.float64)


def _ref_trt(populations, op, om):
    """The reference's vectorised check of its numba collision (test_lbm_collision.py:21-48)."""
    density = np.sum(populations, axis=2)
    momentum = np.einsum("yxd,dc->yxc", populations, C)
    velocity = momentum / density[:, :, None]
    proj = np.einsum("dc,yxc->yxd", C, velocity)
    s2 = np.sum(velocity * velocity, axis=2)
    eq = density[:, :, None] * L.W[None, None, :] * (1.0 + 3.0 * proj + 4.5 * proj * proj - 1.5 * s2[:, :, None])
    delta = populations - eq
    return density, populations - 0.5 * (op + om) * delta - 0.5 * (op - om) * delta[:, :, L.OPP]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_trt_collision_matches_reference_restatement(dtype):
    rng = np.random.default_rng(17)
    f = np.asarray(0.05 + rng.random((5, 7, 9)), dtype=dtype)
    rho, post = L.trt_collision(f, 1.7, 0.4)
    erho, epost = _ref_trt(f.astype(np.float64), 1.7, 0.4)
    np.testing.assert_allclose(rho, erho, rtol=2e-6)
    np.testing.assert_allclose(post, epost, rtol=2e-5, atol=2e-6)


def test_trt_collision_decays_symmetric_and_antisymmetric_parts():
    """TRT definition: post - feq = (1-w+)(f+ - feq+) + (1-w-)(f- - feq-)."""
    rng = np.random.default_rng(3)
    f = 1.0 / 9.0 + 0.01 * rng.random((4, 6, 9))
    op, om = 1.3, 0.7
    rho, post = L.trt_collision(f, op, om)
    r, u = L.macroscopic(f)
    feq = L.equilibrium(r, u)
    neq = f - feq
    sym = 0.5 * (neq + neq[..., L.OPP])
    anti = 0.5 * (neq - neq[..., L.OPP])
    expect = feq + (1 - op) * sym + (1 - om) * anti
    np.testing.assert_allclose(post, expect, atol=1e-14)
    # conservation of mass and momentum
    np.testing.assert_allclose(post.sum(-1), f.sum(-1), rtol=1e-14)
    np.testing.assert_allclose(post @ C, f @ C, atol=1e-15)


def test_equilibrium_moment_round_trip():
    # tests/unit/test_lbm_moments.py:10-32
    rng = np.random.default_rng(305)
    rho = 0.9 + 0.2 * rng.random((3, 5))
    vel = 0.03 * rng.normal(size=(3, 5, 2))
    eq = L.equilibrium(rho, vel)
    r, u = L.macroscopic(eq)
    np.testing.assert_allclose(r, rho, rtol=1e-13)
    np.testing.assert_allclose(u, vel, rtol=1e-13, atol=1e-13)
    # an equilibrium is a fixed point of the collision
    _, post = L.trt_collision(eq, 1.5, 0.6)
    np.testing.assert_allclose(post, eq, atol=1e-15)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_stream_without_solids_is_periodic_roll(dtype):
    # tests/unit/test_lbm_collision.py:52-76
    rng = np.random.default_rng(23)
    post = np.asarray(rng.random((5, 7, 9)), dtype=dtype)
    rho = np.asarray(rng.random((5, 7)) + 1.0, dtype=dtype)
    solid = np.zeros((5, 7), dtype=bool)
    sdf = np.ones((5, 7))
    out = L.moving_wall_stream(post, rho, solid, sdf, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 0.0, 1.0)
    exp = np.empty_like(post)
    for d in range(9):
        exp[..., d] = np.roll(post[..., d], shift=(int(C[d, 1]), int(C[d, 0])), axis=(0, 1))
    np.testing.assert_array_equal(out, exp)


def _scalar_moving_wall(post, density, solid, sdf, omega, scale, pivot):
    """Per-link loop restating kern/lbm.py:263-308 on a unit grid (tiny inputs only)."""
    ny, nx, _ = post.shape
    out = np.zeros_like(post)
    out[..., 0] = post[..., 0]
    for y in range(ny):
        for x in range(nx):
            if solid[y, x]:
                continue
            for d in range(1, 9):
                cx, cy = int(C[d, 0]), int(C[d, 1])
                tx, ty = (x + cx) % nx, (y + cy) % ny
                if not solid[ty, tx]:
                    out[ty, tx, d] += post[y, x, d]
                    continue
                q = min(max(sdf[y, x] / max(sdf[y, x] - sdf[ty, tx], 1e-12), 0.05), 1.0)
                o = int(L.OPP[d])
                if q < 0.5:
                    r = 2 * q * post[y, x, d] + (1 - 2 * q) * post[(y - cy) % ny, (x - cx) % nx, d]
                else:
                    r = post[y, x, d] / (2 * q) + (2 * q - 1) * post[y, x, o] / (2 * q)
                wx = (x + 0.5) + q * cx
                wy = (y + 0.5) + q * cy
                r -= 6 * L.W[d] * density[y, x] * (cx * (-omega * (wy - pivot[1])) + cy * omega * (wx - pivot[0])) * scale
                out[y, x, o] = r
    return out


def test_moving_wall_stream_matches_link_loop_and_reference_check():
    # tests/unit/test_lbm_collision.py:79-132 setup (disc of radius 1.35)
    rng = np.random.default_rng(29)
    post = rng.random((6, 8, 9))
    rho = rng.random((6, 8)) + 1.0
    yy, xx = np.mgrid[:6, :8]
    sdf = np.hypot(xx + 0.5 - 3.5, yy + 0.5 - 2.5) - 1.35
    solid = sdf < 0.0
    out = L.moving_wall_stream(post, rho, solid, sdf, 0.0, 0.0, 1.0, 1.0, 3.0, 2.0, 0.17, 0.08)
    ref = _scalar_moving_wall(post, rho, solid, sdf, 0.17, 0.08, (3.0, 2.0))
    np.testing.assert_allclose(out, ref, rtol=1e-13, atol=1e-13)


def test_halfway_bounce_back_special_case():
    """q = 1/2 everywhere: the reflected population is the outgoing one (no interpolation)."""
    rng = np.random.default_rng(5)
    post = rng.random((5, 6, 9))
    rho = np.ones((5, 6))
    solid = np.zeros((5, 6), bool)
    solid[2, 3] = True
    sdf = np.where(solid, -1.0, 1.0)  # q = 1 / (1 - (-1)) = 0.5
    out = L.moving_wall_stream(post, rho, solid, sdf, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 0.0, 1.0)
    for d in range(1, 9):
        sx, sy = 3 - int(C[d, 0]), 2 - int(C[d, 1])  # fluid neighbour pointing at the solid
        assert out[sy, sx, L.OPP[d]] == post[sy, sx, d]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("side", ["left", "bottom", "top"])
def test_zou_he_imposes_target_velocity(side, dtype):
    """Zou-He closes the unknown populations so that rho*u equals the target exactly
    (up to rounding): a sign or index slip in any of the three formulas fails this."""
    rng = np.random.default_rng(11)
    f = np.asarray(1.0 / 9.0 + 0.01 * rng.random((6, 7, 9)), dtype=dtype)
    ux, uy = 0.04, -0.013
    getattr(L, f"{side}_velocity")(f, ux, uy)
    sl = {"left": (slice(None), 0), "bottom": (0, slice(None)), "top": (-1, slice(None))}[side]
    b = f[sl].astype(np.float64)
    rho = b.sum(-1)
    u = (b @ C) / rho[:, None]
    tol = 1e-14 if dtype == np.float64 else 2e-6
    np.testing.assert_allclose(u[:, 0], ux, atol=tol)
    np.testing.assert_allclose(u[:, 1], uy, atol=tol)


def test_sponge_relaxes_towards_equilibrium():
    rng = np.random.default_rng(31)
    f = rng.random((5, 7, 9))
    eq = rng.random((5, 7, 9))
    s = rng.random((5, 7)) * 0.2
    s[0, 0] = 0.0
    out = L.sponge(f, eq, s)
    np.testing.assert_allclose(out, (1 - s[..., None]) * f + s[..., None] * eq, rtol=1e-15)
    np.testing.assert_array_equal(out[0, 0], f[0, 0])
