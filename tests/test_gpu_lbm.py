"""GPU parity of the D2Q9 TRT LBM against the oracle (fb/solvers/lbm.py, kern/lbm.py).

Bar: the plain forms (FP64, and FP32 storage with FP64 kernel arithmetic) are compared
bit-for-bit with the oracle when both use the same geometry (the GPU's signed distance is
handed to the oracle, since the reference's x**3/x**4 last bits are platform dependent);
with each side computing its own geometry, FP64 agrees to 1e-12 relative. The mixed forms
are compared with the FP64 GPU run within the tolerances stated in each test.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle.lbm import LBMOracle
from tests.helpers import make_scenario

pytestmark = pytest.mark.gpu


def _solver(form="plain"):
    from foilbench_amd.solvers.lbm import LBMSolver

    return LBMSolver(form)


def _foil(sc):
    from foilbench_amd.core.geometry import NacaFoil

    return NacaFoil(sc.foil)


CASES = {
    # static foil, open boundaries: Zou-He rows and inlet, outlet, corners, sponge, Bouzidi
    "static": dict(resolution=(64, 32), controls=((0.0, -8.0),), bounds=((-1.5, 2.5), (-1.0, 1.0)),
                   output_dt=0.02, reynolds=300.0),
    # the same with the foil more than a chord plus eight cells from every edge, run with fused
    # pairs switched on (two steps per launch, the first held in LDS); 96 = 64 + 32 columns leave
    # a partial last tile
    "static-wide": dict(resolution=(96, 64), controls=((0.0, -8.0),), bounds=((-2.5, 3.5), (-2.0, 2.0)),
                        output_dt=0.05, reynolds=300.0),
    # rotating foil: per-substep geometry, uncovered-cell reset, moving-wall term
    "moving": dict(resolution=(64, 32), controls=((0.0, 4.0), (0.2, 30.0)), bounds=((-1.5, 2.5), (-1.0, 1.0)),
                   output_dt=0.02, reynolds=300.0),
    # periodic box, Taylor-Green initial condition, foil outside the domain
    "taylor-green": dict(resolution=(32, 32), initial_condition="taylor-green", foil_in_domain=False,
                         periodic_axes=("x", "y"), bounds=((0.0, 6.283185307179586), (0.0, 6.283185307179586)),
                         output_dt=0.05, controls=((0.0, 0.0),)),
    # channel: periodic x, half-way-style row copies at the walls
    "poiseuille": dict(resolution=(32, 16), initial_condition="poiseuille", foil_in_domain=False,
                       periodic_axes=("x",), bounds=((0.0, 2.0), (-0.5, 0.5)), output_dt=0.05,
                       controls=((0.0, 0.0),)),
}


def _fuse(nx, ny, precision, value):
    """Sets the tuning table's fusion choice for a lattice shape (1 on, 0 off, -1 default)."""
    from foilbench_amd import _native as N

    ctx = N.Context.default()
    prec = N.FB_F64 if precision == "float64" else N.FB_F32
    N.check(ctx.lib.fb_tune_set(N.FB_TUNE_LBM_FUSE, nx, ny, prec, 1, value))


def _run_pair(case, precision, steps, share_geometry):
    sc = make_scenario(precision=precision, **CASES[case])
    if case == "static-wide":
        _fuse(*CASES[case]["resolution"], precision, 1)
    gpu = _solver()
    gpu.initialize(sc, _foil(sc), 0)
    orc = LBMOracle()
    orc.initialize(sc)
    sdf, solid = gpu.geometry_fields()
    np.testing.assert_array_equal(solid, orc.solid)
    if share_geometry:
        orc.sdf = sdf
    f, outlet = gpu.populations()
    np.testing.assert_array_equal(f, orc.f)
    t = 0.0
    reports = []
    for _ in range(steps):
        dt = sc.output_dt
        control = sc.control_at(t + dt)
        rg = gpu.advance(control, dt)
        ro = orc.advance(control, dt)
        t += dt
        reports.append((rg, ro))
        assert rg.substeps == ro["substeps"]
    return gpu, orc, reports


@pytest.mark.parametrize("precision", ["float64", "float32"])
@pytest.mark.parametrize("case", ["static", "static-wide", "taylor-green", "poiseuille"])
def test_lbm_plain_bit_exact_with_shared_geometry(case, precision):
    try:
        gpu, orc, reports = _run_pair(case, precision, steps=4, share_geometry=True)
    finally:
        if case == "static-wide":
            _fuse(*CASES[case]["resolution"], precision, -1)
    f, outlet = gpu.populations()
    np.testing.assert_array_equal(f, orc.f)
    np.testing.assert_array_equal(outlet, orc.outlet)
    rg, ro = reports[-1]
    assert rg.evidence["minimum_population"] == ro["minimum_population"]
    assert rg.max_speed == ro["max_speed"]
    assert rg.evidence["density_excursion"] == ro["density_excursion"]


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_lbm_moving_foil_matches_oracle(precision):
    # Each side recomputes the geometry every substep; the GPU's x**3/x**4 are correctly
    # rounded, NumPy's pow is not, so near-wall values differ at the last-bit level.
    gpu, orc, reports = _run_pair("moving", precision, steps=5, share_geometry=False)
    f, _ = gpu.populations()
    tol = 1e-12 if precision == "float64" else 2e-6
    np.testing.assert_allclose(f, orc.f, rtol=tol, atol=tol * 1e-2)
    sdf, solid = gpu.geometry_fields()
    np.testing.assert_array_equal(solid, orc.solid)
    np.testing.assert_allclose(sdf, orc.sdf, rtol=0, atol=1e-15)
    for rg, ro in reports:
        assert rg.substeps == ro["substeps"]
        assert rg.evidence["maximum_wall_speed"] > 0.0


def test_lbm_export_matches_oracle():
    gpu, orc, _ = _run_pair("static", "float64", steps=3, share_geometry=True)
    st = gpu.export_state()
    vel, rho = orc.export_velocity_density()
    np.testing.assert_array_equal(st.velocity[0], vel)
    np.testing.assert_array_equal(st.density[0], rho)


def test_lbm_rollback_restores_state_bitwise():
    from foilbench_amd.core.models import ControlState, NumericalFailure

    sc = make_scenario(precision="float64", **CASES["static"])
    gpu = _solver()
    gpu.initialize(sc, _foil(sc), 0)
    gpu.advance(sc.control_at(0.02), 0.02)
    before, out_before = gpu.populations()
    rev = gpu.state_revision
    with pytest.raises(NumericalFailure) as err:
        gpu.advance(ControlState(1.0, -8.0, 0.0), 0.02)  # wrong completion time
    assert err.value.reason == "time_contract_failure"
    # forced failure after the lattice steps: an impossible density envelope
    import foilbench_amd.solvers.lbm as mod

    saved = mod._MIN_POPULATION
    mod._MIN_POPULATION = 10.0
    try:
        with pytest.raises(NumericalFailure) as err2:
            gpu.advance(sc.control_at(0.04), 0.02)
        assert err2.value.reason == "invalid_population"
    finally:
        mod._MIN_POPULATION = saved
    after, out_after = gpu.populations()
    np.testing.assert_array_equal(after, before)
    np.testing.assert_array_equal(out_after, out_before)
    assert gpu.state_revision == rev
    # and the next advance equals a fresh run's second advance
    gpu.advance(sc.control_at(0.04), 0.02)
    ref = _solver()
    ref.initialize(sc, _foil(sc), 0)
    ref.advance(sc.control_at(0.02), 0.02)
    ref.advance(sc.control_at(0.04), 0.02)
    np.testing.assert_array_equal(gpu.populations()[0], ref.populations()[0])


@pytest.mark.parametrize("form,tol", [("f32-mixed", 2e-6), ("f16-mixed", 3e-3)])
def test_lbm_mixed_forms_track_fp64(form, tol):
    """Shifted FP32 storage (FP64 collision) and shifted FP16 storage (FP32 collision) against the FP64 run:
    relative L2 error of the lattice velocity over fluid cells after 4 advances."""
    sc = make_scenario(precision="float32", **CASES["static"])
    sc64 = make_scenario(precision="float64", **CASES["static"])
    a = _solver(form)
    a.initialize(sc, _foil(sc), 0)
    b = _solver()
    b.initialize(sc64, _foil(sc64), 0)
    t = 0.0
    for _ in range(4):
        c = sc.control_at(t + sc.output_dt)
        ra = a.advance(c, sc.output_dt)
        rb = b.advance(c, sc.output_dt)
        assert ra.substeps == rb.substeps
        t += sc.output_dt
    _, ua = a._macroscopic()
    _, ub = b._macroscopic()
    _, solid = b.geometry_fields()
    fl = ~solid
    err = np.linalg.norm((ua[fl] - ub[fl]).ravel()) / np.linalg.norm(ub[fl].ravel())
    assert err < tol, err


def _lattice_run(form, nx, ny, channel, fuse, chunks):
    """A harness lattice stepped in advances of the given lengths, with fused pairs switched on, or
    off with two interior rows per thread in the single-step kernel, through the tuning table;
    returns the state and the statistics of each advance."""
    from foilbench_amd import _native as N
    from foilbench_amd.benchmark.lbm_harness import Lattice, LatticeCase
    from foilbench_amd.tuning import _PREC

    ctx = N.Context.default()
    prec = _PREC[form]
    N.check(ctx.lib.fb_tune_set(N.FB_TUNE_LBM_FUSE, nx, ny, prec, 1, 1 if fuse else 0))
    N.check(ctx.lib.fb_tune_set(N.FB_TUNE_LBM_ROWS, nx, ny, prec, 1, 1 if fuse else 2))
    lat = Lattice(LatticeCase(nx=nx, ny=ny, channel=channel), prec, ctx)
    try:
        stats = []
        for n in chunks:
            st = lat.advance(n)
            stats.append((st.min_population, st.max_speed, st.density_excursion, st.nonfinite_populations,
                          st.bad_fluid_density, st.nonfinite_velocity))
        f, outlet = lat.state()
    finally:
        lat.close()
        N.check(ctx.lib.fb_tune_set(N.FB_TUNE_LBM_FUSE, nx, ny, prec, 1, -1))
        N.check(ctx.lib.fb_tune_set(N.FB_TUNE_LBM_ROWS, nx, ny, prec, 1, -1))
    return f, outlet, stats


@pytest.mark.parametrize("form", ["f64", "f32", "f32-mixed", "f16-mixed"])
@pytest.mark.parametrize("nx,ny,channel", [(256, 128, 0), (200, 100, 0), (192, 98, 1)])
def test_lbm_fused_pairs_equal_single_steps(form, nx, ny, channel):
    """Two steps per launch (odd steps with one interior row per thread) against one step per
    launch with two rows per thread: the same bits for the populations, the outlet history and the
    in-kernel postcondition reductions. The shapes leave partial tiles
    (200 x 100, 192 x 98) and include channel walls; advances of odd length end on a single step.
    The harness foil (unit chord at the origin in [-1.6, 4.8] x [-1.6, 1.6]) keeps more than a chord
    plus eight cells from every edge, so every pair of steps is fused when allowed."""
    chunks = (7, 4, 1)
    f1, o1, s1 = _lattice_run(form, nx, ny, channel, True, chunks)
    f0, o0, s0 = _lattice_run(form, nx, ny, channel, False, chunks)
    assert np.isfinite(f1).all()
    np.testing.assert_array_equal(f1, f0)
    np.testing.assert_array_equal(o1, o0)
    assert s1 == s0



def _lattice_run_nt(form, nx, ny, channel, rows, nt, chunks):
    """A harness lattice in single steps with `rows` interior rows per thread and nontemporal population
    traffic on or off (tuning table); returns the state and the statistics of each advance."""
    from foilbench_amd import _native as N
    from foilbench_amd.benchmark.lbm_harness import Lattice, LatticeCase
    from foilbench_amd.tuning import _PREC

    ctx = N.Context.default()
    prec = _PREC[form]
    keys = ((N.FB_TUNE_LBM_FUSE, 0), (N.FB_TUNE_LBM_ROWS, rows), (N.FB_TUNE_LBM_NT, nt))
    for k, v in keys:
        N.check(ctx.lib.fb_tune_set(k, nx, ny, prec, 1, v))
    lat = Lattice(LatticeCase(nx=nx, ny=ny, channel=channel), prec, ctx)
    try:
        stats = []
        for n in chunks:
            st = lat.advance(n)
            stats.append((st.min_population, st.max_speed, st.density_excursion, st.nonfinite_populations,
                          st.bad_fluid_density, st.nonfinite_velocity))
        f, outlet = lat.state()
    finally:
        lat.close()
        for k, _ in keys:
            N.check(ctx.lib.fb_tune_set(k, nx, ny, prec, 1, -1))
    return f, outlet, stats


@pytest.mark.parametrize("form", ["f64", "f32", "f32-mixed", "f16-mixed"])
@pytest.mark.parametrize("rows", [1, 2])
def test_lbm_streaming_population_traffic_changes_no_bit(form, rows):
    """The single-step kernel with nontemporal population loads and stores (the form chosen for lattices
    larger than the caches) against the cached form: the same populations, outlet history and
    postcondition statistics, with one and two interior rows per thread, on a walled box and a channel."""
    for nx, ny, channel in ((200, 100, 0), (192, 98, 1)):
        chunks = (5, 3)
        f1, o1, s1 = _lattice_run_nt(form, nx, ny, channel, rows, 1, chunks)
        f0, o0, s0 = _lattice_run_nt(form, nx, ny, channel, rows, 0, chunks)
        assert np.isfinite(f1).all()
        np.testing.assert_array_equal(f1, f0)
        np.testing.assert_array_equal(o1, o0)
        assert s1 == s0
