"""Pin the diagnostics restatement (oracle/metrics.py) and the host wake spectrum
(foilbench_amd/core/metrics.py): the reference's wake-metrics fixture and unit cases
(tests/unit/test_interpolation_metrics.py), and closed forms that a wrong sign, axis or offset would fail."""

from __future__ import annotations

import json

import numpy as np
import pytest

from oracle import metrics as M
from oracle.mac import Dom
from tests.conftest import GOLDEN


def test_wake_metrics_fixture():
    fx = json.loads((GOLDEN / "wake-metrics.json").read_text())
    dom = Dom(fx["bounds"], *fx["resolution"])
    vel = np.asarray(fx["velocity"], np.float64)
    solid = np.asarray(fx["solid"], bool)
    assert M.wake_width(vel, dom, fx["pivot_x"], fx["chord"], fx["freestream_u"], solid) == pytest.approx(
        fx["expected"]["wake_width"])
    assert M.recirculation_area(vel, dom, fx["pivot_x"], solid) == pytest.approx(fx["expected"]["recirculation_area"])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_energy_enstrophy_and_divergence_closed_forms(dtype):
    dom = Dom(((-1.0, 3.0), (-0.5, 1.5)), 16, 8)
    yc = np.linspace(-0.5 + 0.125, 1.5 - 0.125, 8)
    xc = np.linspace(-1.0 + 0.125, 3.0 - 0.125, 16)
    X, Y = np.meshgrid(xc, yc)
    uniform = np.broadcast_to(np.asarray([1.5, -0.5], dtype), (8, 16, 2)).copy()
    assert M.kinetic_energy(uniform) == pytest.approx(0.5 * (1.5 ** 2 + 0.5 ** 2), rel=1e-6)
    assert M.enstrophy(uniform, dom) == 0.0 and M.divergence_l2(uniform, dom) == 0.0
    # solid-body rotation u = -W y, v = W x: vorticity 2 W everywhere (exact for linear fields, ends included),
    # divergence zero
    w = 0.75
    rot = np.stack((-w * Y, w * X), axis=-1).astype(dtype)
    np.testing.assert_allclose(M.vorticity(rot, dom), 2.0 * w, rtol=1e-5)
    assert M.enstrophy(rot, dom) == pytest.approx(2.0 * w * w, rel=1e-5)
    assert M.divergence_l2(rot, dom) == pytest.approx(0.0, abs=1e-5)
    # u = a x, v = b y: divergence a + b, vorticity zero
    lin = np.stack((0.5 * X, -2.0 * Y), axis=-1).astype(dtype)
    assert M.divergence_l2(lin, dom) == pytest.approx(1.5, rel=1e-5)
    assert M.enstrophy(lin, dom) == pytest.approx(0.0, abs=1e-8)


def test_wake_rows_and_recirculating_cells_counted_by_hand():
    dom = Dom(((0.0, 8.0), (0.0, 4.0)), 8, 4)  # centres x = 0.5 .. 7.5
    vel = np.ones((4, 8, 2))
    vel[1, 3, 0] = 0.85   # deficit 0.15 > 0.1 but x = 3.5 is not beyond pivot + chord = 4
    vel[2, 5, 0] = 0.85   # counted
    vel[3, 6, 0] = 0.95   # deficit 0.05: not counted
    vel[0, 7, 0] = 0.5    # counted, unless solid
    assert M.wake_width(vel, dom, 3.0, 1.0, 1.0) == 2.0
    solid = np.zeros((4, 8), bool)
    solid[0, 7] = True
    assert M.wake_width(vel, dom, 3.0, 1.0, 1.0, solid) == 1.0
    vel[1, 1, 0] = -0.1   # upstream of the pivot x = 3
    vel[2, 4, 0] = -0.2
    vel[3, 7, 0] = -0.3
    assert M.recirculation_area(vel, dom, 3.0) == 2.0
    assert M.recirculation_area(vel, dom, 3.0, solid) == 2.0
    solid[3, 7] = True
    assert M.recirculation_area(vel, dom, 3.0, solid) == 1.0


def test_bilinear_sampling_closed_forms():
    # tests/unit/test_interpolation_metrics.py:22-29, and exact reproduction of a linear field
    dom = Dom(((0.0, 2.0), (0.0, 1.0)), 8, 4)
    const = np.empty((4, 8, 2))
    const[...] = (2.0, -0.5)
    pts = np.asarray([[0.1, 0.1], [1.0, 0.5], [1.9, 0.9]])
    np.testing.assert_allclose(M.sample_vector(const, pts, dom), [[2.0, -0.5]] * 3)
    xc = (np.arange(8) + 0.5) * 0.25
    yc = (np.arange(4) + 0.5) * 0.25
    X, Y = np.meshgrid(xc, yc)
    lin = np.stack((1.0 + 2.0 * X - Y, 3.0 * Y), axis=-1)
    inner = np.asarray([[0.2, 0.2], [1.0, 0.6], [1.8, 0.8], [0.13, 0.77]])
    want = np.stack((1.0 + 2.0 * inner[:, 0] - inner[:, 1], 3.0 * inner[:, 1]), axis=1)
    np.testing.assert_allclose(M.sample_vector(lin, inner, dom), want, atol=1e-14)
    # outside the centres the coordinate is clipped: the edge value
    edge = M.sample_vector(lin, np.asarray([[-5.0, 0.5]]), dom)
    np.testing.assert_allclose(edge, M.sample_vector(lin, np.asarray([[0.125, 0.5]]), dom), atol=1e-15)
    # periodic: one period away samples the same point
    pdom = Dom(((0.0, 2.0), (0.0, 1.0)), 8, 4, ("x", "y"))
    a = M.sample_vector(lin, np.asarray([[0.07, 0.31]]), pdom)
    b = M.sample_vector(lin, np.asarray([[2.07, 1.31]]), pdom)
    np.testing.assert_allclose(a, b, atol=1e-12)
    assert M.sample_vector(lin.astype(np.float32), inner, dom).dtype == np.float32


def _link_case(periodic=()):
    dom = Dom(((0.0, 5.0), (0.0, 5.0)), 5, 5, periodic)
    c = np.linspace(0.5, 4.5, 5)
    X, Y = np.meshgrid(c, c)
    centres = np.stack((X, Y), axis=-1)
    solid = np.zeros((5, 5), bool)
    return dom, centres, solid


def test_cut_link_speed_hand_computed():
    dom, centres, solid = _link_case()
    solid[2, 2] = True
    # signed distance +0.5 in fluid and -0.5 in the solid cell: every wall point sits halfway along its link
    sdf = np.full((5, 5), 0.5)
    sdf[2, 2] = -0.5
    vel = np.zeros((5, 5, 2))
    vel[1, 1] = (0.6, 0.6)   # lower-left neighbour, link (+1, +1): |0.6 + 0.6| / sqrt 2
    assert M.cut_link_adjacent_normal_speed(vel, solid, sdf, centres, dom, (0.0, 0.0), 0.0) == pytest.approx(1.2 / np.sqrt(2))
    vel[2, 1] = (1.0, 0.0)   # left neighbour, link (+1, 0): |1 * 1| / 1 = 1, the larger
    assert M.cut_link_adjacent_normal_speed(vel, solid, sdf, centres, dom, (0.0, 0.0), 0.0) == 1.0
    vel[2, 1] = (0.0, 1.0)   # tangential to its link: only the diagonal one counts
    assert M.cut_link_adjacent_normal_speed(vel, solid, sdf, centres, dom, (0.0, 0.0), 0.0) == pytest.approx(1.2 / np.sqrt(2))
    # at rest, rotating at W: the wall point of each link sits halfway, and the projection of W x (p - pivot) on
    # the link. About the solid cell's centre (2.5, 2.5) every r is along or across its link: all zero.
    vel[...] = 0.0
    w = 0.4
    assert M.cut_link_adjacent_normal_speed(vel, solid, sdf, centres, dom, (2.5, 2.5), w) == pytest.approx(0.0, abs=1e-15)
    # about the origin: the axis links see 2.5 W (e.g. from cell (1, 2) along +x: p = (2.0, 2.5), wall velocity
    # (-2.5 W, 2 W)), the diagonals from (3, 1) along (-1, +1) and from (1, 3) along (+1, -1) see
    # |(2 W, -3 W) . (-1, 1)| / sqrt 2 = 5 W / sqrt 2, the largest
    assert M.cut_link_adjacent_normal_speed(vel, solid, sdf, centres, dom, (0.0, 0.0), w) == pytest.approx(5 * w / np.sqrt(2))
    # the along-link component of a rigid rotation does not depend on where the wall point sits: a fraction
    # clipped to 0.05 gives the same speed
    sdf[2, 2] = -100.0
    assert M.cut_link_adjacent_normal_speed(vel, solid, sdf, centres, dom, (0.0, 0.0), w) == pytest.approx(5 * w / np.sqrt(2))


def test_cut_link_wrapped_links_only_on_periodic_axes():
    for periodic, want in (((), 0.0), (("x",), 1.0)):
        dom, centres, solid = _link_case(periodic)
        solid[2, 0] = True
        sdf = np.full((5, 5), 0.5)
        sdf[2, 0] = -0.5
        vel = np.zeros((5, 5, 2))
        vel[2, 4] = (1.0, 0.0)   # its +x link wraps onto the solid cell at column 0
        got = M.cut_link_adjacent_normal_speed(vel, solid, sdf, centres, dom, (0.0, 0.0), 0.0)
        assert got == want, (periodic, got)


def test_particle_census_hand_placed():
    dom = Dom(((0.0, 4.0), (0.0, 2.0)), 4, 2)
    solid = np.zeros((2, 4), bool)
    solid[0, 0] = True
    # cell (0, 1): 5 particles, (0, 2): 1, (1, 3): 4 (one beyond the box, clipped in), the rest empty; 2 in
    # the solid cell do not count
    pts = [(1.5, 0.5)] * 5 + [(2.2, 0.1)] + [(3.5, 1.5)] * 3 + [(9.0, 9.0)] + [(0.5, 0.5)] * 2
    c = M.particle_census(np.asarray(pts), solid, dom)
    assert c["empty_fluid_cell_fraction"] == 4 / 7
    assert c["underfilled_fluid_cell_fraction"] == 5 / 7
    assert c["max_particles_per_fluid_cell"] == 5.0
    assert c["p05_particles_per_fluid_cell"] == 0.0
    assert c["p95_particles_per_fluid_cell"] == pytest.approx(float(np.percentile([0, 0, 0, 0, 1, 4, 5], 95)))


def test_wake_spectrum_unit_cases():
    # tests/unit/test_interpolation_metrics.py:40-65 on the host spectrum the benchmark runner uses
    from foilbench_amd.core.metrics import analyze_wake_probe

    dt = 0.01
    t = np.arange(400) * dt
    s = analyze_wake_probe(0.2 + 0.7 * np.sin(2.0 * np.pi * 2.5 * t), dt, chord=1.0, freestream_speed=2.0)
    assert s.sample_count == 400 and s.frequency_resolution == pytest.approx(0.25)
    assert s.transverse_rms == pytest.approx(0.7 / np.sqrt(2.0), rel=0.01)
    assert s.dominant_frequency == pytest.approx(2.5) and s.strouhal_number == pytest.approx(1.25)
    assert s.dominant_power_fraction > 0.6
    s = analyze_wake_probe(np.ones(32), 0.1, chord=1.0, freestream_speed=1.0)
    assert (s.transverse_rms, s.dominant_frequency, s.strouhal_number, s.dominant_power_fraction) == (0, 0, 0, 0)
    with pytest.raises(ValueError):
        analyze_wake_probe(np.ones(7), 0.1, 1.0, 1.0)
