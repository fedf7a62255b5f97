"""Pin the oracle against the reference's own golden vectors (spec/conformance/*, copied
under tests/golden/) and the values its conformance tests assert."""

from __future__ import annotations

import json

import numpy as np
import pytest

from oracle.geometry import NacaFoil, cell_centers
from oracle.lbm import sponge_strength_profile
from oracle.rng import PCG32


def _load(golden, name):
    return json.loads((golden / name).read_text())


def test_pcg32_matches_shared_vectors(golden):
    # spec/conformance/pcg32.json; tests/conformance/test_shared_fixtures.py:18-33
    doc = _load(golden, "pcg32.json")
    assert len(doc["cases"]) == 3
    for case in doc["cases"]:
        rng = PCG32(case["seed"], stream=case["stream"])
        assert [rng.next_uint32() for _ in case["uint32"]] == case["uint32"]
        vals = PCG32(case["seed"], stream=case["stream"]).random((len(case["uint32"]),))
        assert [f"0x{int(v):08x}" for v in vals.view(np.uint32)] == case["float32_bits"]


def test_pcg32_default_stream_vector():
    # tests/unit/test_rng.py:6-21
    rng = PCG32(42, stream=54)
    assert [rng.next_uint32() for _ in range(8)] == [
        2707161783, 2068313097, 3122475824, 2211639955, 3215226955, 3421331566, 3217466285, 2167406445]
    v = PCG32(7).random((3, 5))
    assert v.shape == (3, 5) and v.dtype == np.float32 and np.all((v >= 0) & (v < 1))


def test_naca2412_vectors(golden):
    # spec/conformance/naca2412.json; tests/conformance/test_shared_fixtures.py:36-82
    doc = _load(golden, "naca2412.json")
    tol = doc["absolute_tolerances"]
    foil = NacaFoil(doc["foil"]["naca"], doc["foil"]["chord"], tuple(doc["foil"]["pivot"]))
    up, lo = foil.surfaces(np.asarray(doc["surface_x"], dtype=np.float64))
    np.testing.assert_allclose(up, doc["surface_upper"], atol=tol["surface"])
    np.testing.assert_allclose(lo, doc["surface_lower"], atol=tol["surface"])
    for q in doc["queries"]:
        pts = np.asarray(q["points"], dtype=np.float64)
        d = foil.signed_distance(pts, q["angle_degrees"])
        np.testing.assert_allclose(d, q["signed_distance"], atol=tol["signed_distance"])
        np.testing.assert_allclose(foil.normals(pts, q["angle_degrees"]), q["normals"], atol=tol["normal"])
        np.testing.assert_array_equal(d <= 0.0, q["contains"])


def test_geometry_v1_vectors(golden):
    # spec/conformance/geometry-v1.json (NACA 0012, chord 2.5, 23 degrees)
    doc = _load(golden, "geometry-v1.json")
    tol = doc["absolute_tolerances"]
    desc = doc["descriptor"]
    foil = NacaFoil(desc["naca"], desc["chord"], tuple(desc["pivot"]))
    up, lo = foil.surfaces(np.asarray(doc["surface_x"]))
    np.testing.assert_allclose(up, doc["surface_upper"], atol=tol["surface"])
    np.testing.assert_allclose(lo, doc["surface_lower"], atol=tol["surface"])
    pts = np.asarray(doc["points"], dtype=np.float64)
    np.testing.assert_allclose(foil.signed_distance(pts, doc["angle_degrees"]), doc["signed_distance"],
                               atol=tol["signed_distance"])
    np.testing.assert_allclose(foil.normals(pts, doc["angle_degrees"]), doc["normals"], atol=tol["normal"])
    np.testing.assert_array_equal(foil.contains(pts, doc["angle_degrees"]), doc["contains"])
    np.testing.assert_allclose(foil.wall_velocity(pts, doc["angular_velocity_degrees"]), doc["wall_velocity"],
                               atol=tol["wall_velocity"])
    assert abs(foil.maximum_radius - doc["maximum_radius"]) <= tol["radius"]


def test_lbm_sponge_fixture(golden):
    # spec/conformance/lbm-boundary.json + tests/conformance/test_revision5_conformance.py:187-211
    sp = _load(golden, "lbm-boundary.json")["sponge"]
    prof = sponge_strength_profile(160, 96, False, False, False)
    assert prof[0, 80] == pytest.approx(sp["transverse_maximum"])
    assert prof[48, 159] == pytest.approx(sp["outlet_maximum"])
    assert prof[0, 159] == pytest.approx(max(sp["transverse_maximum"], sp["outlet_maximum"]))
    assert sponge_strength_profile(160, 96, False, False, True)[0, 80] == 0.0
    # base width max(3, floor(min(nx, ny)/16)): zero strength at and beyond the width
    w = max(3, 96 // 16)
    assert prof[w, 80] == 0.0 and prof[w - 1, 80] > 0.0


def test_cell_centers_are_linspace_midpoints():
    c = cell_centers(((-1.5, 2.5), (-1.0, 1.0)), 8, 4)
    assert c.shape == (4, 8, 2)
    np.testing.assert_allclose(c[0, :, 0], -1.5 + 0.25 + 0.5 * np.arange(8), atol=1e-15)
    assert c[-1, -1, 0] == 2.5 - 0.25 and c[-1, -1, 1] == 1.0 - 0.25
