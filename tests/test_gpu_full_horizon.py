"""GPU parity at the configurations' horizons.

- The 1024 x 512 headline lattice (NACA 2412 at AoA 12, u = 0.05, Re 1000): 1000 lattice steps in FP32 storage
  and 500 in FP64 against the C restatement of the oracle (oracle/lbm_oracle.c, itself bit-equal to
  oracle/lbm.py kernel_steps), from the same initial populations with the GPU's signed distance: identical
  populations and outlet history, with the autotuner's launch shape and with fused pairs.
- Served lattice requests of the serving mix at their full step counts (the first request of each grid shape
  up to 512 x 256: 3731, 672 and 1799 steps): the plain FP64 lattice equals the C oracle bit for bit, and the
  served form (shifted FP32 storage, FP64 collision, batched launches) is within 1e-5 relative of the oracle in
  velocity and density deviation.
- Served Stable Fluids requests at their full step counts (the first of each grid shape up to 512 x 256: 3693,
  2004 and 245 steps): within 1e-5 relative of the plain FP64 form per face field; and the plain form of the
  first of them against the FP64 oracle (oracle/mac.py) over its first 60 steps at 1e-9.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _oracle_args(case, hd, sdf, solid):
    from oracle import lbm as L

    dx = (case.bounds[0][1] - case.bounds[0][0]) / case.nx
    dy = (case.bounds[1][1] - case.bounds[1][0]) / case.ny
    _, op, om = case.relaxation()
    u = case.u_lattice
    target = tuple(np.asarray([1.0, 0.0], dtype=hd) * u)
    feqb = L.equilibrium(np.ones((1, 1), hd), np.asarray(target, hd).reshape(1, 1, 2))[0, 0]
    sp = L.sponge_strength_profile(case.nx, case.ny, False, False, False)
    return (solid, sdf, op, om, u, target, feqb, sp), (case.bounds, dx, dy, case.pivot, 0.0, u, False, False, False)


@pytest.mark.parametrize("prec_name,steps,fuse", [("f32", 1000, None), ("f32", 1000, 1), ("f64", 500, None)])
def test_headline_lattice_long_horizon_bit_exact(prec_name, steps, fuse):
    from foilbench_amd import _native as N
    from foilbench_amd.benchmark.lbm_harness import Lattice, LatticeCase
    from foilbench_amd.tuning import _PREC, Autotuner
    from oracle import lbm_c

    ctx = N.Context.default()
    case = LatticeCase()
    prec = _PREC[prec_name]
    if fuse is None:
        Autotuner(ctx).lbm(case.nx, case.ny, prec_name)
    else:
        N.check(ctx.lib.fb_tune_set(N.FB_TUNE_LBM_FUSE, case.nx, case.ny, prec, 1, fuse))
    lat = Lattice(case, prec, ctx)
    try:
        f0, out0 = lat.state()
        sdf = np.empty((case.ny, case.nx))
        solid = np.empty((case.ny, case.nx), np.uint8)
        N.check(ctx.lib.fb_lbm_get_geometry(lat.handle, N.ptr(sdf), N.ptr(solid)))
        lat.advance(steps)
        f_gpu, out_gpu = lat.state()
    finally:
        lat.close()
        for param in (N.FB_TUNE_LBM_FUSE, N.FB_TUNE_LBM_ROWS, N.FB_TUNE_LBM_LIST_BLOCKS):
            N.check(ctx.lib.fb_tune_set(param, case.nx, case.ny, prec, 1, -1))
    a, b = _oracle_args(case, f0.dtype, sdf, solid.astype(bool))
    f, outlet = lbm_c.kernel_steps(f0, *a, out0.reshape(case.ny, 1, 9), *b, steps)
    print(f"{prec_name} 1024x512 {steps} steps fuse={fuse}: max |gpu - oracle| = {np.max(np.abs(f_gpu - f)):.3g}")
    assert np.isfinite(f_gpu).all()
    np.testing.assert_array_equal(f_gpu, f)
    np.testing.assert_array_equal(out_gpu.reshape(case.ny, 9), outlet.reshape(case.ny, 9))


def _first_per_shape(solver, shapes):
    from foilbench_amd.serving.mix import serving_mix

    out = {}
    for r in serving_mix(256):
        if r.solver == solver and (r.nx, r.ny) in shapes and (r.nx, r.ny) not in out:
            out[(r.nx, r.ny)] = r
    return [out[s] for s in shapes]


SHAPES = ((128, 64), (256, 128), (512, 256))


def test_served_lattices_at_full_horizon_match_the_oracle():
    from foilbench_amd import _native as N
    from foilbench_amd.serving.engine import Server, _build, _close
    from foilbench_amd.serving.mix import LBM
    from oracle import lbm as L
    from oracle import lbm_c

    reqs = _first_per_shape(LBM, SHAPES)
    ctx = N.Context.default()
    srv = Server(sf_workers=2, lbm_chunk=32)
    try:
        served = {o.id: o for o in srv.run(reqs, keep_fields=True, realtime=False)["outcomes"]}
    finally:
        srv.close()
    for r in reqs:
        print(f"lattice request {r.id}: {r.nx}x{r.ny}, {r.steps} steps", flush=True)
        assert served[r.id].status == 0 and served[r.id].steps == r.steps
        s, handle, params, _ = _build(r, "plain", ctx)
        try:
            f0, out0 = s.populations()
            sdf, solid = s.geometry_fields()
            dom = s._scenario.domain
            vs = s._reference_speed / s._lattice_speed
            feqb = np.asarray([params.feq_boundary[d] for d in range(9)])
            sp = L.sponge_strength_profile(dom.nx, dom.ny, False, False, False)
            f, outlet = lbm_c.kernel_steps(f0, solid, sdf, params.omega_plus, params.omega_minus, params.u_lattice,
                                           (params.target_ux, params.target_uy), feqb, sp, out0, dom.bounds,
                                           dom.dx, dom.dy, s._scenario.foil.pivot, 0.0, params.wall_scale, False,
                                           False, False, r.steps)
            # the plain FP64 lattice, one launch per step: the oracle's bits
            N.check(ctx.lib.fb_lbm_run_fixed(handle, C.byref(params), r.steps))
            ctx.synchronize()
            fg, og = s.populations()
            np.testing.assert_array_equal(fg, f)
            np.testing.assert_array_equal(og.reshape(-1, 9), outlet.reshape(-1, 9))
        finally:
            _close(s)
        rho, vel = L.macroscopic(f)
        vel = vel * vs
        got = served[r.id].fields
        ev = _rel(got["vel"], vel)
        er = _rel(got["rho"] - 1.0, rho - 1.0)
        print(f"  served vs oracle: velocity {ev:.2e}, density deviation {er:.2e}", flush=True)
        assert ev <= 1e-5 and er <= 1e-5, (r.id, ev, er)


def test_served_stable_fluids_at_full_horizon():
    from foilbench_amd import _native as N
    from foilbench_amd.serving.engine import Server, _build, _close, run_plain
    from foilbench_amd.serving.mix import SF
    from oracle.mac import StableFluidsOracle

    reqs = _first_per_shape(SF, SHAPES)
    srv = Server(sf_workers=4, lbm_chunk=32)
    try:
        served = {o.id: o for o in srv.run(reqs, keep_fields=True, realtime=False)["outcomes"]}
    finally:
        srv.close()
    plain = {o.id: o for o in run_plain(reqs, keep_fields=True)["outcomes"]}
    for r in reqs:
        so, po = served[r.id], plain[r.id]
        assert so.status == 0 and so.sf_status == 0 and so.steps == r.steps
        errs = {k: _rel(so.fields[k], po.fields[k]) for k in ("u", "v")}
        print(f"stable-fluids request {r.id}: {r.nx}x{r.ny}, {r.steps} steps, served vs plain {errs}", flush=True)
        assert max(errs.values()) <= 1e-5, (r.id, errs)
    # the plain form against the FP64 oracle over the first steps of the first request
    r = reqs[0]
    ctx = N.Context.default()
    s, handle, params, pose = _build(r, "plain", ctx)
    try:
        sc = s._scenario
        orc = StableFluidsOracle()
        orc.initialize(sc)
        t = 0.0
        for _ in range(60):
            t += sc.output_dt
            c = sc.control_at(t)
            s.advance(c, sc.output_dt)
            orc.advance(c, sc.output_dt)
        u, v = s.faces()
    finally:
        _close(s)
    assert _rel(u, orc.u) < 1e-9 and _rel(v, orc.v) < 1e-9, (_rel(u, orc.u), _rel(v, orc.v))
