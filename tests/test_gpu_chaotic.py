"""The chaotic-wake scenario (scenarios/airfoil/chaotic-experimental.json: NACA 2412 held at 35 degrees, Re 10^4,
160 x 96, FP32, skew-symmetric RK2 Stable Fluids transport with stable_cfl 0.4) on the GPU solvers.

Each solver runs its first advances beside the oracle (FP32 storage: 2e-5 relative for Stable Fluids and
PIC/FLIP fields, the lattice bit for bit with a shared signed distance), then runs on alone over a longer horizon,
where the separated wake must stay finite, keep its postconditions and grow unsteady (the transverse velocity one
chord behind the foil varies)."""

from __future__ import annotations

import numpy as np
import pytest

from tests import fidelity as F

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("solver_id", F.SOLVER_IDS)
def test_chaotic_scenario_matches_oracle_then_stays_unsteady(solver_id):
    sc = F.scenario_file("scenarios/airfoil/chaotic-experimental.json")
    assert sc.solver_options["stable_advection"] == "skew-rk2"
    gpu = F.GpuRun(solver_id)
    gpu.initialize(sc, sc.seed)
    orc = F.OracleRun(solver_id)
    orc.initialize(sc, sc.seed)
    if solver_id == "lbm-d2q9":
        sdf, solid = gpu.s.geometry_fields()
        np.testing.assert_array_equal(solid, orc.o.solid)
        orc.o.sdf = sdf
    t = 0.0
    for _ in range(3):
        t += sc.output_dt
        c = sc.control_at(t)
        eg = gpu.advance(c, sc.output_dt)
        eo = orc.advance(c, sc.output_dt)
        assert gpu.report.substeps == eo["substeps"]
    if solver_id == "lbm-d2q9":
        f, _ = gpu.s.populations()
        np.testing.assert_array_equal(f, orc.o.f)
    else:
        err = _rel(gpu.velocity(), orc.velocity())
        assert err < 2e-5, (solver_id, err)
    probe = []
    # probe 1.5 chords behind the pivot, on its row (the runner's wake probe point)
    py = int((sc.foil.pivot[1] - sc.domain.bounds[1][0]) / sc.domain.dy)
    px = int((sc.foil.pivot[0] + 1.5 * sc.foil.chord - sc.domain.bounds[0][0]) / sc.domain.dx)
    for _ in range(120):
        t += sc.output_dt
        gpu.advance(sc.control_at(t), sc.output_dt)
        probe.append(float(gpu.velocity()[py, px, 1]))
    assert np.isfinite(gpu.velocity()).all()
    assert np.ptp(np.asarray(probe)) > 1e-3, (solver_id, np.ptp(np.asarray(probe)))
    diag = gpu.s.diagnostics().values
    assert all(np.isfinite(v) for v in diag.values())
