"""GPU parity at the benchmark configurations' own sizes.

- C1 Stable Fluids FP64, NACA 0012 at AoA 8 on 256 x 128, Re 500, CG to 1e-6: three advances against the oracle
  at 1e-9, equal substep counts, CG and Jacobi counts within one.
- C2 D2Q9 TRT on the 1024 x 512 harness lattice (NACA 2412 at AoA 12, u = 0.05, Re 1000, Lambda = 3/16):
  lattice steps against oracle/lbm.py kernel_steps bit for bit in FP64 and FP32 storage (shared signed distance),
  single and fused-pair launches.
- Masks at every configuration's size, computed by each side on its own, bit-equal: C1 256 x 128 (0012 at -8),
  C2 1024 x 512 (2412 at -12), C3 2048 x 1024 (0012 at -15), C4 1024 x 512 (4412 at -10).
- C3 Stable Fluids at 2048 x 1024 (NACA 0012 at AoA 15, Re 2000, CG to 1e-8): one advance of the mixed
  (FP32 sweeps, FP64 refinement) and precond32 (FP64 CG, FP32 DST-I preconditioner) forms against the FP64
  oracle, 1e-6 relative per face field.
- C4 PIC/FLIP seeding at 8 particles per fluid cell on 1024 x 512 (NACA 4412 at AoA 10, PCG32 seed 42) bit-equal
  to the oracle's, and one FP64 advance at 8 per cell on 256 x 128 against the oracle at 1e-9.
"""

from __future__ import annotations

import functools

import numpy as np
import pytest

from foilbench_amd.benchmark.legs import pic_scenario, sf_scenario

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _foil(sc):
    from foilbench_amd.core.geometry import NacaFoil

    return NacaFoil(sc.foil)


# ------------------------------------------------------------------------------------------ masks

MASK_CASES = {
    "C1": (256, 128, ((-1.5, 2.5), (-1.0, 1.0)), "0012", -8.0),
    "C2": (1024, 512, ((-1.6, 4.8), (-1.6, 1.6)), "2412", -12.0),
    "C3": (2048, 1024, ((-1.5, 2.5), (-1.0, 1.0)), "0012", -15.0),
    "C4": (1024, 512, ((-1.6, 4.8), (-1.6, 1.6)), "4412", -10.0),
}


@pytest.mark.parametrize("config", list(MASK_CASES))
def test_masks_at_config_size_are_bit_equal(config):
    from foilbench_amd.core.geometry import NacaFoil
    from foilbench_amd.core.models import DomainSpec, FoilSpec
    from oracle.geometry import NacaFoil as OracleFoil
    from oracle.geometry import cell_centers

    nx, ny, bounds, naca, angle = MASK_CASES[config]
    dom = DomainSpec(2, bounds, (nx, ny))
    sdf, solid = NacaFoil(FoilSpec(naca, 1.0, (0.0, 0.0))).cell_sdf(dom, angle)
    orc = OracleFoil(naca, 1.0, (0.0, 0.0))
    pts = cell_centers(bounds, nx, ny).reshape(-1, 2)
    osdf = orc.signed_distance(pts, angle).reshape(ny, nx)
    print(f"{config}: {int(solid.sum())} solid cells, max |sdf - oracle| = {np.max(np.abs(sdf - osdf)):.2e}")
    np.testing.assert_array_equal(solid, osdf <= 0.0)
    assert solid.sum() > 100
    np.testing.assert_allclose(sdf, osdf, rtol=0, atol=1e-14)


# ------------------------------------------------------------------------------------------ C1

def test_c1_stable_fluids_fp64_256x128_matches_oracle():
    from foilbench_amd.solvers.stable_fluids import StableFluidsSolver
    from oracle.mac import StableFluidsOracle

    sc = sf_scenario(256, 128, -8.0, 500.0, "float64", 1e-6)
    gpu = StableFluidsSolver("plain")
    gpu.initialize(sc, _foil(sc), 0)
    orc = StableFluidsOracle()
    orc.initialize(sc)
    np.testing.assert_array_equal(gpu.solid_mask(), orc.solid)
    t = 0.0
    for _ in range(3):
        t += sc.output_dt
        c = sc.control_at(t)
        rg = gpu.advance(c, sc.output_dt)
        ro = orc.advance(c, sc.output_dt)
        assert rg.substeps == ro["substeps"] == 1
        assert abs(rg.evidence["pressure_iterations"] - ro["pressure_iterations"]) <= 1
        assert abs(rg.evidence["viscosity_iterations"] - ro["viscosity_iterations"]) <= 1
    u, v = gpu.faces()
    assert _rel(u, orc.u) < 1e-9 and _rel(v, orc.v) < 1e-9, (_rel(u, orc.u), _rel(v, orc.v))
    assert _rel(gpu.pressure(), orc.pressure) < 1e-8


# ------------------------------------------------------------------------------------------ C2

def _harness(prec_name, fuse):
    from foilbench_amd import _native as N
    from foilbench_amd.benchmark.lbm_harness import Lattice, LatticeCase
    from foilbench_amd.tuning import _PREC

    ctx = N.Context.default()
    case = LatticeCase()
    prec = _PREC[prec_name]
    N.check(ctx.lib.fb_tune_set(N.FB_TUNE_LBM_FUSE, case.nx, case.ny, prec, 1, 1 if fuse else 0))
    return ctx, case, Lattice(case, prec, ctx), prec


@pytest.mark.parametrize("prec_name,fuse", [("f64", False), ("f32", False), ("f32", True)])
def test_c2_lattice_1024x512_bit_exact_with_oracle(prec_name, fuse):
    """Six lattice steps of the headline lattice on the GPU and in oracle/lbm.py kernel_steps, from the same
    initial populations and the GPU's signed distance: identical populations and outlet history."""
    from foilbench_amd import _native as N
    from oracle import lbm as L

    ctx, case, lat, prec = _harness(prec_name, fuse)
    try:
        f0, out0 = lat.state()
        sdf = np.empty((case.ny, case.nx))
        solid = np.empty((case.ny, case.nx), np.uint8)
        N.check(ctx.lib.fb_lbm_get_geometry(lat.handle, N.ptr(sdf), N.ptr(solid)))
        solid = solid.astype(bool)
        steps = 6
        lat.advance(steps)
        f_gpu, out_gpu = lat.state()
    finally:
        lat.close()
        N.check(ctx.lib.fb_tune_set(N.FB_TUNE_LBM_FUSE, case.nx, case.ny, prec, 1, -1))
    hd = f0.dtype
    dx = (case.bounds[0][1] - case.bounds[0][0]) / case.nx
    dy = (case.bounds[1][1] - case.bounds[1][0]) / case.ny
    _, op, om = case.relaxation()
    u = case.u_lattice
    target = tuple(np.asarray([1.0, 0.0], dtype=hd) * u)
    feqb = L.equilibrium(np.ones((1, 1), hd), np.asarray(target, hd).reshape(1, 1, 2))[0, 0]
    sp = L.sponge_strength_profile(case.nx, case.ny, False, False, False)
    f = f0.copy()
    outlet = out0.reshape(case.ny, 1, 9).copy()
    f, outlet = L.kernel_steps(f, solid, sdf, op, om, u, target, feqb, sp, outlet, case.bounds, dx, dy, case.pivot,
                               0.0, u, False, False, False, steps)
    assert np.isfinite(f_gpu).all()
    np.testing.assert_array_equal(f_gpu, f)
    np.testing.assert_array_equal(out_gpu.reshape(case.ny, 9), outlet.reshape(case.ny, 9))


# ------------------------------------------------------------------------------------------ C3

@functools.lru_cache(maxsize=1)
def _c3_oracle():
    """One FP64 oracle advance of the C3 workload (about 40 s of NumPy/SciPy), shared by both forms."""
    from oracle.mac import StableFluidsOracle

    sc = sf_scenario(2048, 1024, -15.0, 2000.0, "float64", 1e-8)
    orc = StableFluidsOracle()
    orc.initialize(sc)
    ro = orc.advance(sc.control_at(sc.output_dt), sc.output_dt)
    return sc, orc, ro


@pytest.mark.parametrize("form", ["mixed", "precond32"])
def test_c3_reduced_precision_2048x1024_against_fp64_oracle(form):
    from foilbench_amd.solvers.stable_fluids import StableFluidsSolver

    sc, orc, ro = _c3_oracle()
    gpu = StableFluidsSolver(form)
    gpu.initialize(sc, _foil(sc), 0)
    np.testing.assert_array_equal(gpu.solid_mask(), orc.solid)
    rg = gpu.advance(sc.control_at(sc.output_dt), sc.output_dt)
    u, v = gpu.faces()
    p = gpu.pressure()
    gpu.close()
    assert rg.substeps == ro["substeps"]
    assert rg.evidence["pressure_relative_residual"] <= 1e-8 * (1 + 1e-6)
    errs = {"u": _rel(u, orc.u), "v": _rel(v, orc.v), "p": _rel(p, orc.pressure)}
    print(f"C3 {form}: CG {rg.evidence['pressure_iterations']} (oracle {ro['pressure_iterations']}), errors {errs}")
    assert errs["u"] < 1e-6 and errs["v"] < 1e-6, errs
    assert errs["p"] < 1e-5, errs


# ------------------------------------------------------------------------------------------ C4

def test_c4_seeding_8_per_cell_1024x512_bit_exact():
    from foilbench_amd.solvers.pic_flip import PicFlipSolver
    from oracle.geometry import NacaFoil as OracleFoil
    from oracle.mac import Dom
    from oracle.pic import seed_particles
    from oracle.rng import PCG32

    sc = pic_scenario(1024, 512, "float32")
    gpu = PicFlipSolver("plain", particles_per_cell=8)
    gpu.initialize(sc, _foil(sc), 42)
    pos, _ = gpu.particles()
    solid = gpu.solid_mask()
    dom = Dom.of(sc.domain)
    orc_solid = OracleFoil("4412", 1.0, (0.0, 0.0)).mask(dom.bounds, dom.nx, dom.ny, -10.0)
    np.testing.assert_array_equal(solid, orc_solid)
    want = seed_particles(orc_solid, dom, PCG32(42, stream=71), np.float32, per_cell=8)
    assert pos.shape == want.shape == (8 * int((~orc_solid).sum()), 2)
    np.testing.assert_array_equal(pos, want)


def test_c4_pic_fp64_8_per_cell_advance_matches_oracle():
    from foilbench_amd.solvers.pic_flip import PicFlipSolver
    from oracle.pic import PicFlipOracle

    sc = pic_scenario(256, 128, "float64")
    gpu = PicFlipSolver("plain", particles_per_cell=8)
    gpu.initialize(sc, _foil(sc), 42)
    orc = PicFlipOracle()
    orc.initialize(sc, seed=42, per_cell=8)
    pos, vel = gpu.particles()
    np.testing.assert_array_equal(pos, orc.positions)
    c = sc.control_at(sc.output_dt)
    rg = gpu.advance(c, sc.output_dt)
    ro = orc.advance(c, sc.output_dt)
    assert rg.substeps == ro["substeps"]
    pos, vel = gpu.particles()
    assert _rel(pos, orc.positions) < 1e-9 and _rel(vel, orc.velocities) < 1e-9
