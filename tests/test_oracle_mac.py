"""Pin the Stable Fluids / MAC oracle (oracle/mac.py) with closed forms, a dense direct
solve and the reference's own projection tests (tests/unit/test_projection.py)."""

from __future__ import annotations

import json

import numpy as np
import pytest

from oracle import mac as M


def _dense(fluid, dx, dy, periodic=()):
    _, apply = M.masked_operator(fluid, dx, dy, periodic)
    n = fluid.size
    A = np.empty((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = apply(e)
    return A


def test_mac_boundary_fixture(golden):
    # spec/conformance/mac-boundary.json + tests/conformance/test_revision5_conformance.py:169-185
    doc = json.loads((golden / "mac-boundary.json").read_text())
    assert doc["periodic_duplicate"] == "endpoint-average"
    dom = M.Dom(((0.0, 2.0), (-1.0, 1.0)), 8, 6)
    u = np.full((6, 9), -3.0)
    v = np.full((7, 8), -4.0)
    M.apply_domain_boundaries(u, v, dom, (1.25, -0.5))
    np.testing.assert_allclose(u[:, 0], 1.25)
    np.testing.assert_allclose(u[0, :], 1.25)
    np.testing.assert_allclose(u[-1, :], 1.25)
    np.testing.assert_allclose(v[0, :], -0.5)
    np.testing.assert_allclose(v[-1, :], -0.5)
    # zero-gradient outlet (prescribed-inlet-zero-gradient-outlet)
    np.testing.assert_array_equal(u[1:-1, -1], u[1:-1, -2])
    # periodic endpoint average
    pdom = M.Dom(((0.0, 2.0), (-1.0, 1.0)), 8, 6, ("x",))
    u = np.arange(54, dtype=float).reshape(6, 9)
    v = np.zeros((7, 8))
    M.apply_domain_boundaries(u, v, pdom, (0.0, 0.0))
    np.testing.assert_array_equal(u[:, 0], u[:, -1])
    # interior rows: the endpoint average; the y-edge rows then take the freestream
    np.testing.assert_allclose(u[1:-1, 0], 0.5 * (np.arange(1, 5) * 9 + np.arange(1, 5) * 9 + 8))
    assert u[0, 0] == 0.0 and u[-1, 0] == 0.0


def test_operator_is_symmetric_positive_definite_and_matches_dense_solve():
    fluid = np.ones((6, 8), dtype=bool)
    fluid[2:4, 3:5] = False
    A = _dense(fluid, 0.25, 0.3)
    np.testing.assert_allclose(A, A.T, atol=1e-12)
    assert np.linalg.eigvalsh(A).min() > 0.0
    rng = np.random.default_rng(4)
    b = rng.normal(size=fluid.shape)
    b[~fluid] = 0.0
    p, info, its, rel = M.solve_masked_poisson(b, fluid, 0.25, 0.3, (), 1e-12, 200)
    assert info == 0 and rel <= 1e-12 * (1 + 1e-6)
    np.testing.assert_allclose(p.ravel(), np.linalg.solve(A, b.ravel()), rtol=1e-9, atol=1e-12)
    # the operator is the 5-point negative Laplacian with p = 0 ghosts at the box edge
    i = (3, 1)
    k = i[0] * 8 + i[1]
    assert A[k, k] == pytest.approx(2 / 0.25**2 + 2 / 0.3**2)
    assert A[k, k + 1] == pytest.approx(-1 / 0.25**2)
    assert A[k, k + 8] == pytest.approx(-1 / 0.3**2)


def test_dst_preconditioner_is_exact_inverse_on_a_solid_free_box():
    fluid = np.ones((9, 12), dtype=bool)
    A = _dense(fluid, 0.1, 0.2)
    pre = M.dst_preconditioner(fluid, 0.1, 0.2)
    rng = np.random.default_rng(8)
    r = rng.normal(size=fluid.size)
    np.testing.assert_allclose(A @ pre(r), r, rtol=1e-12, atol=1e-12)
    b = rng.normal(size=fluid.shape)
    p, info, its, rel = M.solve_masked_poisson(b, fluid, 0.1, 0.2, (), 1e-10, 50)
    assert its == 1 and info == 0 and rel < 1e-12


def test_zero_rhs_returns_zero_without_iterations():
    fluid = np.ones((5, 6), dtype=bool)
    p, info, its, rel = M.solve_masked_poisson(np.zeros((5, 6)), fluid, 0.1, 0.1)
    assert info == 0 and its == 0 and rel == 0.0 and not p.any()


def test_masked_projection_reduces_fluid_divergence():
    # tests/unit/test_projection.py:32-58
    dom = M.Dom(((-2.0, 2.0), (-1.0, 1.0)), 64, 32)
    rng = np.random.default_rng(7)
    vel = rng.normal(0.0, 0.1, (32, 64, 2))
    vel[:, :, 0] += 1.0
    solid = np.zeros((32, 64), dtype=bool)
    solid[12:20, 28:36] = True
    wall = np.zeros_like(vel)
    u, v = M.cell_to_faces(vel)
    M.apply_domain_boundaries(u, v, dom, (1.0, 0.0))
    M.enforce_solid_faces(u, v, solid, wall)
    before = (u[:, 1:] - u[:, :-1]) / dom.dx + (v[1:] - v[:-1]) / dom.dy
    u2, v2, rep, _ = M.project_faces(u, v, dom, solid, wall, (1.0, 0.0), 0.01)
    after = (u2[:, 1:] - u2[:, :-1]) / dom.dx + (v2[1:] - v2[:-1]) / dom.dy
    assert rep[2]
    assert np.linalg.norm(after[~solid]) < 0.25 * np.linalg.norm(before[~solid])


def test_periodic_face_diffusion_matches_discrete_fourier_decay():
    # tests/unit/test_projection.py:90-112
    dom = M.Dom(((0.0, 2.0), (0.0, 1.0)), 32, 16, ("x", "y"))
    phase = 2 * np.pi * np.arange(33) / 32
    u = np.broadcast_to(np.sin(phase), (16, 33)).copy()
    nu, dt = 0.2, 0.03
    out, its, res, ok = M.implicit_diffuse_scalar(u, nu, dt, dom, 1e-10)
    lam = 4.0 * np.sin(np.pi / 32) ** 2 / dom.dx**2
    assert ok
    np.testing.assert_allclose(out[:, :-1], u[:, :-1] / (1 + nu * dt * lam), atol=1e-8)
    np.testing.assert_allclose(out[:, 0], out[:, -1], atol=1e-12)


def test_rk2_backtrace_rigid_rotation_closed_form():
    # tests/unit/test_projection.py:146-160
    dom = M.Dom(((-2.0, 2.0), (-2.0, 2.0)), 64, 64)
    x = np.linspace(-2 + 0.5 * dom.dx, 2 - 0.5 * dom.dx, 64)
    xx, yy = np.meshgrid(x, x)
    vel = np.stack((-yy, xx), axis=2)
    dep = M.rk2_backtrace(vel, np.asarray([[1.0, 0.0], [0.0, 1.0]]), 0.1, dom)
    np.testing.assert_allclose(dep, [[0.995, -0.1], [0.1, 0.995]], atol=1e-12)


def test_limited_maccormack_stays_inside_local_bounds_and_preserves_constants():
    # tests/unit/test_projection.py:163-184
    dom = M.Dom(((0.0, 1.0), (0.0, 1.0)), 32, 32, ("x", "y"))
    vel = np.zeros((32, 32, 2))
    vel[:, :16, 0] = 0.8
    vel[10:21, :, 1] = -0.4
    lo, hi = M.local_velocity_bounds(vel, dom)
    adv = M.advect_velocity(vel, 0.02, dom, True)
    assert np.all(adv >= lo - 1e-12) and np.all(adv <= hi + 1e-12)
    const = np.zeros((16, 24, 2))
    const[..., 0] = 1.25
    const[..., 1] = -0.2
    d2 = M.Dom(((-2.0, 2.0), (-1.0, 1.0)), 24, 16)
    np.testing.assert_allclose(M.advect_velocity(const, 0.03, d2, True), const, rtol=1e-15)


# ----------------------------------------------------------------------------- face-native transports

def test_face_advection_keeps_parallel_shear_flows_with_the_right_face_offsets():
    """u = a y (v = 0) departs horizontally and samples u at its own height, so every u face
    keeps its value; likewise v = b x (u = 0) on v faces. A wrong sampling offset shifts the
    sampled height (or abscissa) by half a cell and changes the field."""
    dom = M.Dom(((-1.0, 1.0), (-0.5, 0.5)), 32, 16)
    pu, pv = M.face_positions(dom)
    u = (0.3 * pu[:, 1]).reshape(16, 33)
    v = np.zeros((17, 32))
    for mac in (False, True):
        au, av = M.advect_faces(u, v, 0.01, dom, mac)
        np.testing.assert_allclose(au, u, rtol=0, atol=1e-15)
        np.testing.assert_array_equal(av, v)
    u = np.zeros((16, 33))
    v = (0.2 * pv[:, 0]).reshape(17, 32)
    for mac in (False, True):
        au, av = M.advect_faces(u, v, 0.01, dom, mac)
        np.testing.assert_array_equal(au, u)
        np.testing.assert_allclose(av, v, rtol=0, atol=1e-15)


def test_face_advection_translates_a_linear_field_exactly_in_the_interior():
    """A uniform stream (U, V) carries u = c + a x + b y, v = const: the departure point of a face
    is its position minus (U, V) dt, and bilinear sampling is exact for a linear field, so away
    from the clamped edges first_u = c + a (x - U dt) + b (y - V dt)."""
    dom = M.Dom(((0.0, 2.0), (0.0, 1.0)), 40, 20)
    pu, _ = M.face_positions(dom)
    U, V, dt = 0.5, -0.25, 0.02
    u = (U + 0.0 * pu[:, 0]).reshape(20, 41)
    v = np.full((21, 40), V)
    # the velocity field is uniform, so the face values themselves must stay (U, V)
    au, av = M.advect_faces(u, v, dt, dom, True)
    np.testing.assert_allclose(au, U, atol=1e-15)
    np.testing.assert_allclose(av, V, atol=1e-15)
    vel = M.faces_to_cell(u, v)
    f = (0.7 + 1.3 * pu[:, 0] - 0.4 * pu[:, 1]).reshape(20, 41)
    first = M._face_departure_sample(f, vel, pu, (0.0, 0.5), dt, dom)
    want = 0.7 + 1.3 * (pu[:, 0] - U * dt) - 0.4 * (pu[:, 1] - V * dt)
    want = want.reshape(20, 41)
    np.testing.assert_allclose(first[2:-2, 2:-2], want[2:-2, 2:-2], rtol=0, atol=1e-13)


def test_face_local_bounds_replicate_edges():
    f = np.arange(12.0).reshape(3, 4)
    lo, hi = M.face_local_bounds(f)
    assert lo[0, 0] == 0.0 and hi[0, 0] == 5.0 and lo[2, 3] == 6.0 and hi[2, 3] == 11.0
    assert lo[1, 1] == 0.0 and hi[1, 1] == 10.0


def test_centered_derivative_closed_forms():
    """Periodic central differences of sin(kx) equal cos(kx) sin(kh)/h exactly (the duplicated
    endpoint repeats the first derivative); second-order one-sided edges make numpy.gradient
    exact on quadratics."""
    n, L = 32, 2.0 * np.pi
    h = L / n
    x = np.arange(n + 1) * h
    f = np.tile(np.sin(3.0 * x), (4, 1))
    d = M.centered_derivative(f, h, 1, True, duplicate_endpoint=True)
    want = np.cos(3.0 * x) * np.sin(3.0 * h) / h
    np.testing.assert_allclose(d[:, :-1], np.tile(want[:-1], (4, 1)), atol=1e-12)
    np.testing.assert_array_equal(d[:, -1], d[:, 0])
    g = np.sin(3.0 * x[:-1])[:, None] * np.ones((1, 3))
    dg = M.centered_derivative(g, h, 0, True)
    np.testing.assert_allclose(dg[:, 0], np.cos(3.0 * x[:-1]) * np.sin(3.0 * h) / h, atol=1e-12)
    y = np.linspace(0.0, 1.0, 9)
    q = np.tile(2.0 * y**2 - y + 0.5, (3, 1))
    np.testing.assert_allclose(M.centered_derivative(q, y[1] - y[0], 1, False), np.tile(4.0 * y - 1.0, (3, 1)),
                               atol=1e-12)


def test_cross_velocity_and_skew_rate_of_a_uniform_stream():
    dom = M.Dom(((0.0, 2.0), (0.0, 1.0)), 16, 8)
    u = np.full((8, 17), 0.75)
    v = np.full((9, 16), -0.5)
    vu, uv = M.cross_velocity_on_faces(u, v, dom)
    np.testing.assert_array_equal(vu, np.full_like(u, -0.5))
    np.testing.assert_array_equal(uv, np.full_like(v, 0.75))
    assert M.skew_face_advection_rate(u, v, dom) == pytest.approx(0.75 / dom.dx + 0.5 / dom.dy, rel=1e-15)
    ku, kv = M.skew_symmetric_convection(u, v, dom)
    np.testing.assert_allclose(ku, 0.0, atol=1e-12)
    np.testing.assert_allclose(kv, 0.0, atol=1e-12)


def test_skew_symmetric_convection_conserves_kinetic_energy_on_a_periodic_box():
    """With central differences and wrap-around, sum(u * (u du/dx + d(uu)/dx) / 2) and its
    cross-velocity counterpart telescope to zero for each component, so the convection neither
    creates nor destroys discrete kinetic energy (over the logical faces). A dropped term or a
    wrong sign breaks the cancellation."""
    dom = M.Dom(((0.0, 1.0), (0.0, 1.0)), 24, 20, ("x", "y"))
    rng = np.random.default_rng(5)
    u = rng.normal(0.0, 1.0, (20, 25))
    v = rng.normal(0.0, 1.0, (21, 24))
    u[:, -1] = u[:, 0]
    v[-1, :] = v[0, :]
    ku, kv = M.skew_symmetric_convection(u, v, dom)
    eu = float(np.sum(u[:, :-1] * ku[:, :-1]))
    ev = float(np.sum(v[:-1, :] * kv[:-1, :]))
    scale = float(np.sum(np.abs(u[:, :-1] * ku[:, :-1])) + np.sum(np.abs(v[:-1, :] * kv[:-1, :])))
    assert abs(eu) < 1e-12 * scale and abs(ev) < 1e-12 * scale
    # the purely advective half alone does not conserve it
    vu, _ = M.cross_velocity_on_faces(u, v, dom)
    adv = u * M.centered_derivative(u, dom.dx, 1, True, True) + vu * M.centered_derivative(u, dom.dy, 0, True)
    assert abs(float(np.sum(u[:, :-1] * adv[:, :-1]))) > 1e-6 * scale


def test_skew_rk2_keeps_a_uniform_stream_and_imposes_the_boundaries():
    dom = M.Dom(((-1.0, 1.0), (-0.5, 0.5)), 16, 8)
    u = np.full((8, 17), 1.0)
    v = np.zeros((9, 16))
    solid = np.zeros((8, 16), bool)
    wall = np.zeros((8, 16, 2))
    au, av = M.advect_faces_skew_rk2(u, v, 0.01, dom, solid, wall, (1.0, 0.0))
    np.testing.assert_allclose(au, 1.0, atol=1e-14)
    np.testing.assert_allclose(av, 0.0, atol=1e-14)
    solid[3:5, 7:9] = True
    au, av = M.advect_faces_skew_rk2(u, v, 0.01, dom, solid, wall, (1.0, 0.0))
    assert np.all(au[3:5, 7:10] == 0.0) and np.all(av[3:6, 7:9] == 0.0)
