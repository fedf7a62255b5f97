"""Request loop over the native server (fb_server_*): continuous batching on one GPU.

Requests arrive on the wall clock (their ``arrival_s``), wait for admission, and are admitted
while their estimated device memory fits the budget and the number in flight stays under
``max_inflight``. Admission builds the request's device state through the solver lifecycle
(``initialize`` on the default context) and hands the handle to the server, which runs Stable
Fluids requests on its worker streams and lattice requests in its continuous batch. Results are
polled as they finish, their fields read back if asked for, and their memory freed, which admits
the next waiting requests.

Forms:
* ``serving``: Stable Fluids in FP64 storage with the drop-in's FP64 pressure CG (the reference's zero
  start and stopping rule), pressure solves of same-shape requests batched, so every served Stable Fluids
  request equals its plain result bit for bit; lattices in FP32 storage shifted by the lattice weights with
  FP64 collision arithmetic, batched. (FP32 faces, a warm-started CG, or the FP32 DST-I preconditioner put
  the fields of long requests beyond 1e-5 of the plain form: 1.5e-3 and 1.1e-5 after 3693 steps of a
  128 x 64 request at AoA 15.7 for the first two; 2.0e-5 after 1811 steps of a 256 x 128 request at
  Re 6588 and AoA 5.9 for the preconditioner, whose solves differ within the CG tolerance and whose
  unsteady wakes amplify the difference.)
* ``plain``: the drop-in form, FP64, one scenario at a time on the default stream with no server
  (substep calls for Stable Fluids, one launch per lattice step).
"""

from __future__ import annotations

import ctypes as C
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from .. import _native as N
from ..core.geometry import NacaFoil, pose_struct
from ..solvers.lbm import LBMSolver
from ..solvers.stable_fluids import StableFluidsSolver
from .mix import LBM, SF, Request

# Device bytes per cell of a request's state, from the allocations of fb_mac_create/Poisson set-up
# and fb_lbm_create, with a margin; plus a fixed allowance per request.
_SF_BYTES_PER_CELL = 430
_LBM_BYTES_PER_CELL = {"f64": 750, "f32-mixed": 500, "f16-mixed": 400, "f32": 500}
_FIXED_BYTES = 8 << 20
# Initial guess of the serving form's pressure solves (fb_sf_params.warm_pressure): 0, zero (the
# reference's); 1, the previous substep's pressure; 2, its linear extrapolation from the two previous ones.
SF_WARM_PRESSURE = 0
# Storage precision of the serving form's Stable Fluids faces.
SF_PRECISION = "float64"
# Pressure solve of the serving form's Stable Fluids requests (a StableFluidsSolver form): "plain", the
# drop-in's FP64 CG; "precond32", FP64 CG with the DST-I preconditioner sweeps in FP32 (measured over the
# full-horizon mix: 14% more Stable Fluids throughput, but two of 124 requests 1.1e-5 and 2.0e-5 from FP64).
SF_SOLVER_FORM = "plain"


def estimate_bytes(req: Request, form: str) -> int:
    if req.solver == SF:
        return _FIXED_BYTES + _SF_BYTES_PER_CELL * req.cells
    lbm_form = "f64" if form == "plain" else "f32-mixed"
    return _FIXED_BYTES + _LBM_BYTES_PER_CELL[lbm_form] * req.cells


@dataclass
class Outcome:
    id: int
    solver: str
    cells: int
    steps: int
    status: int
    sf_status: int
    arrival_ms: float
    admitted_ms: float
    started_ms: float
    finished_ms: float
    pressure_iterations: int = 0
    fields: dict = field(default_factory=dict)
    error: str = ""  # why a request failed before reaching the server (its build raised)

    @property
    def latency_ms(self) -> float:
        return self.finished_ms - self.arrival_ms


def _build(req: Request, form: str, ctx: N.Context):
    """Device state of a request through the solver lifecycle; returns (solver, handle, params, pose)."""
    if req.solver == SF:
        sc = req.scenario("float64" if form == "plain" else SF_PRECISION)
        # serving form: SF_SOLVER_FORM's pressure solve, batched across same-shape requests by the server
        if form == "plain":
            s = StableFluidsSolver("plain", ctx=ctx)
        else:
            s = StableFluidsSolver(SF_SOLVER_FORM, ctx=ctx, warm_pressure=SF_WARM_PRESSURE)
        s.initialize(sc, NacaFoil(sc.foil), 0)
        params = s._params(sc.output_dt, sc.control_at(0.0))
        return s, s._handle, params, pose_struct(-req.aoa_degrees)
    sc = req.scenario("float64" if form == "plain" else "float32")
    s = LBMSolver("f64" if form == "plain" else "f32-mixed", ctx=ctx)
    s.initialize(sc, NacaFoil(sc.foil), 0)
    return s, s._handle, s._params(0.0), None


def _fields(req: Request, solver) -> dict:
    if req.solver == SF:
        u, v = solver.faces()
        return {"u": np.asarray(u, dtype=np.float64), "v": np.asarray(v, dtype=np.float64)}
    rho, vel = solver._macroscopic()
    return {"rho": np.asarray(rho, dtype=np.float64), "vel": np.asarray(vel, dtype=np.float64)}


def _close(solver) -> None:
    if isinstance(solver, StableFluidsSolver):
        solver.close()
    else:
        lib = solver._ctx.lib
        if solver._handle is not None:
            lib.fb_lbm_destroy(solver._handle)
            solver._handle = None


class Server:
    """The native request loop (fb_server) with memory-based admission in front of it."""

    def __init__(self, sf_workers: int = 8, lbm_chunk: int = 32, use_graphs: bool = True,
                 memory_budget: int | None = None, max_inflight: int = 64, sf_batch: int | None = None,
                 sf_gather_us: float = 2000.0, builders: int = 4, ctx: N.Context | None = None) -> None:
        self.ctx = ctx or N.Context.default()
        self.lib = self.ctx.lib
        self.builders = max(1, int(builders))
        self.tuning: dict = {}
        self._local = threading.local()
        self._builder_ctxs: list[N.Context] = []
        sf_batch = sf_workers if sf_batch is None else sf_batch
        opt = N.fb_server_options(sf_workers, lbm_chunk, 1 if use_graphs else 0, sf_batch, sf_gather_us)
        h = C.c_void_p()
        N.check(self.lib.fb_server_create(self.ctx.device, C.byref(opt), C.byref(h)))
        self.handle = h
        if memory_budget is None:
            free = C.c_int64()
            total = C.c_int64()
            N.check(self.lib.fb_mem_info(self.ctx.device, C.byref(free), C.byref(total)))
            memory_budget = int(0.8 * free.value)
        self.memory_budget = int(memory_budget)
        self.max_inflight = int(max_inflight)
        self.options = {"sf_workers": sf_workers, "lbm_chunk": lbm_chunk, "use_graphs": bool(use_graphs),
                        "sf_batch": sf_batch, "sf_gather_us": sf_gather_us, "builders": self.builders,
                        "memory_budget": self.memory_budget, "max_inflight": self.max_inflight}

    @classmethod
    def tuned(cls, requests: list[Request], sf_workers: int = 16, tuner=None, **kw) -> "Server":
        """A server whose launch shapes and lattice chunk were picked by the autotuner for the grid
        shapes of `requests` (the choices are in `server.tuning`)."""
        from ..tuning import Autotuner

        tuner = tuner or Autotuner(kw.get("ctx"))
        choices = tuner.for_mix(requests, kw.get("sf_batch") or sf_workers)
        srv = cls(sf_workers=sf_workers, lbm_chunk=int(choices["lbm_chunk"]), **kw)
        srv.tuning = choices
        srv.options["tuning"] = choices
        return srv

    def clock_ms(self) -> float:
        ms = C.c_double()
        N.check(self.lib.fb_server_clock(self.handle, C.byref(ms)))
        return ms.value

    def submit(self, req: Request, handle, params, pose) -> None:
        if req.solver == SF:
            N.check(self.lib.fb_server_submit_sf(self.handle, req.id, handle, C.byref(params), C.byref(pose),
                                                 req.steps))
        else:
            N.check(self.lib.fb_server_submit_lbm(self.handle, req.id, handle, C.byref(params), req.steps))

    def poll(self, max_results: int = 64) -> list[N.fb_server_result]:
        buf = (N.fb_server_result * max_results)()
        n = C.c_int()
        N.check(self.lib.fb_server_poll(self.handle, buf, max_results, C.byref(n)))
        return [buf[k] for k in range(n.value)]

    def close(self) -> None:
        if self.handle is not None:
            N.check(self.lib.fb_server_destroy(self.handle))
            self.handle = None

    def __del__(self) -> None:  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def _builder_ctx(self) -> N.Context:
        c = getattr(self._local, "ctx", None)
        if c is None:
            c = N.Context(self.ctx.device)
            self._local.ctx = c
            self._builder_ctxs.append(c)
        return c

    def _build_and_submit(self, r: Request):
        t0 = self.clock_ms()
        solver, handle, params, pose = _build(r, "serving", self._builder_ctx())
        self.submit(r, handle, params, pose)
        return solver, self.clock_ms() - t0

    def run(self, requests: list[Request], keep_fields=False, realtime: bool = True) -> dict:
        """Serve `requests` (arrivals on the wall clock when `realtime`, else all at once). `keep_fields`:
        True, or a set of request ids whose final fields are read back. Admitted requests are built
        (solver initialisation) by `builders` threads, each with its own context, and handed to the
        server as soon as they are ready."""
        order = sorted(requests, key=lambda r: (r.arrival_s, r.id))
        by_id = {r.id: r for r in order}
        c0 = self.clock_ms()
        arrival = {r.id: (c0 + 1e3 * r.arrival_s if realtime else c0) for r in order}
        waiting: list[Request] = []
        inflight: dict[int, list] = {}  # id -> [future, need, t_admitted]
        used = 0
        outcomes: dict[int, Outcome] = {}
        nxt = 0
        peak_used = 0
        build_ms = 0.0
        pool = ThreadPoolExecutor(max_workers=self.builders)
        try:
            while nxt < len(order) or waiting or inflight:
                now = self.clock_ms()
                while nxt < len(order) and arrival[order[nxt].id] <= now:
                    waiting.append(order[nxt])
                    nxt += 1
                progressed = False
                # admission in arrival order, by memory and slots
                while waiting and len(inflight) < self.max_inflight:
                    r = waiting[0]
                    need = estimate_bytes(r, "serving")
                    if used + need > self.memory_budget and inflight:
                        break
                    waiting.pop(0)
                    inflight[r.id] = [pool.submit(self._build_and_submit, r), need, self.clock_ms()]
                    used += need
                    peak_used = max(peak_used, used)
                    progressed = True
                # a build that failed never reaches the server: report it here
                for rid, (fut, need, t_adm) in list(inflight.items()):
                    if fut.done() and fut.exception() is not None:
                        r = by_id[rid]
                        inflight.pop(rid)
                        used -= need
                        t = self.clock_ms()
                        outcomes[rid] = Outcome(r.id, r.solver, r.cells, 0, -1, 0, arrival[rid], t_adm,
                                                t, t, error=repr(fut.exception()))
                        progressed = True
                for res in self.poll():
                    r = by_id[res.id]
                    fut, need, t_adm = inflight.pop(res.id)
                    solver, bms = fut.result()
                    build_ms += bms
                    out = Outcome(r.id, r.solver, r.cells, int(res.steps_done), int(res.status), int(res.sf_status),
                                  arrival[r.id], t_adm, res.started_ms, res.finished_ms, int(res.pressure_iterations))
                    if _keep(keep_fields, r.id) and res.status == 0:
                        out.fields = _fields(r, solver)
                    _close(solver)
                    used -= need
                    outcomes[r.id] = out
                    progressed = True
                if not progressed:
                    time.sleep(0.0005)
        finally:
            pool.shutdown(wait=True)
        out = summarize([outcomes[r.id] for r in order], self.options)
        out["host_build_s"] = build_ms / 1e3
        out["peak_admitted_bytes"] = peak_used
        return out


def _keep(keep_fields, rid: int) -> bool:
    return bool(keep_fields) if isinstance(keep_fields, bool) else rid in keep_fields


def run_plain(requests: list[Request], keep_fields=False, ctx: N.Context | None = None) -> dict:
    """The drop-in form, one request after another with no overlap (FP64; substep calls for Stable
    Fluids, one launch per lattice step)."""
    ctx = ctx or N.Context.default()
    lib = ctx.lib
    outcomes = []
    t_start = time.perf_counter()
    for r in sorted(requests, key=lambda q: (q.arrival_s, q.id)):
        t_adm = (time.perf_counter() - t_start) * 1e3
        solver, handle, params, pose = _build(r, "plain", ctx)
        ctx.synchronize()
        t0 = (time.perf_counter() - t_start) * 1e3
        status, sf_status, its, done = 0, 0, 0, 0
        if r.solver == SF:
            rep = N.fb_sf_report()
            for k in range(r.steps):
                status = lib.fb_sf_substep(handle, C.byref(params), C.byref(pose), C.byref(rep))
                its += int(rep.pressure_iterations)
                if status or rep.status:
                    sf_status = int(rep.status)
                    break
                done = k + 1
        else:
            status = lib.fb_lbm_run_fixed(handle, C.byref(params), r.steps)
            done = r.steps if status == 0 else 0
        ctx.synchronize()
        t1 = (time.perf_counter() - t_start) * 1e3
        out = Outcome(r.id, r.solver, r.cells, done, int(status), sf_status, t_adm, t_adm, t0, t1, its)
        if _keep(keep_fields, r.id) and status == 0:
            out.fields = _fields(r, solver)
        _close(solver)
        outcomes.append(out)
    return summarize(outcomes, {"form": "plain"})


def summarize(outcomes: list[Outcome], options: dict) -> dict:
    ok = [o for o in outcomes if o.status == 0 and o.sf_status == 0]
    t0 = min(o.arrival_ms for o in outcomes)
    t1 = max(o.finished_ms for o in outcomes)
    cell_steps = float(sum(o.cells * o.steps for o in ok))
    lat = np.asarray([o.latency_ms for o in ok]) if ok else np.zeros(1)
    busy = sum(o.finished_ms - o.started_ms for o in ok)
    return {
        "requests": len(outcomes),
        "completed": len(ok),
        "failed": [(o.id, o.status, o.sf_status) for o in outcomes if o not in ok],
        "cell_steps": cell_steps,
        "makespan_s": (t1 - t0) / 1e3,
        "cell_steps_per_s": cell_steps / max((t1 - t0) / 1e3, 1e-9),
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "mean_run_ms": busy / max(len(ok), 1),
        "by_solver": {s: {"count": sum(1 for o in ok if o.solver == s),
                          "cell_steps": float(sum(o.cells * o.steps for o in ok if o.solver == s))}
                      for s in (SF, LBM)},
        "options": options,
        "outcomes": outcomes,
    }
