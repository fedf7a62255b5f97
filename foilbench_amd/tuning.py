"""Autotuner: picks launch shapes and schedules on the device by timing candidates.

Tuned here, per grid shape and precision, and stored in the library's tuning table
(`fb_tune_set`), which the launches consult:
* points per thread and rows per block of the DST-I kernel (the tile of the preconditioner's
  x-direction transform), timed as whole preconditioner applications on stacked grids
  (`fb_poisson_time_precond`);
* for a lattice, whether two steps run per launch (`lbm_pair_kernel`), the interior rows per
  thread and the nontemporal population traffic of the single-step kernel (jointly) and the blocks
  striding over its flagged cells, timed as lattice steps (`fb_lbm_time_fixed`).
For the request loop it also picks the lattice-batch chunk (steps per graph) by timing
batched steps of the mix's lattice shapes (`fb_lbm_run_batch`).

Every tuned choice leaves each result bit unchanged (rows are independent transforms; a lattice
cell is stepped by the same arithmetic whichever block, launch shape or fused pair takes it), so
tuning changes only speed. Results can be saved to and loaded from a JSON file.
"""

from __future__ import annotations

import ctypes as C
import json
import time
from pathlib import Path

from . import _native as N

_PREC = {"f64": N.FB_F64, "f32": N.FB_F32, "f32-mixed": N.FB_F32_MIXED, "f16-mixed": N.FB_F16_MIXED}


class Autotuner:
    def __init__(self, ctx: N.Context | None = None, path: str | Path | None = None, reps: int = 20) -> None:
        self.ctx = ctx or N.Context.default()
        self.lib = self.ctx.lib
        self.reps = reps
        self.path = Path(path) if path else None
        self.choices: dict[str, dict] = {}
        if self.path and self.path.exists():
            self.choices = json.loads(self.path.read_text())
            for rec in self.choices.values():
                if "param" in rec:
                    N.check(self.lib.fb_tune_set(rec["param"], rec["nx"], rec["ny"], rec["precision"], rec["grids"],
                                                 rec["value"]))
                for entry in rec.get("set", []):
                    N.check(self.lib.fb_tune_set(*entry))

    def save(self) -> None:
        if self.path:
            self.path.parent.mkdir(parents=True, exist_ok=True)
            self.path.write_text(json.dumps(self.choices, indent=1, sort_keys=True))

    def _set(self, param, nx, ny, prec, grids, value) -> None:
        N.check(self.lib.fb_tune_set(param, nx, ny, prec, grids, value))

    # ------------------------------------------------------------------ Poisson preconditioner
    def poisson(self, nx: int, ny: int, dx: float, dy: float, precision: int = N.FB_F32, grids: int = 1,
                key_grids: int | None = None) -> dict:
        """Points per thread and rows per DST block for `grids` stacked nx x ny grids; stored under
        batch size `key_grids` (default `grids`; 0 = any batch size). "value" is the rows per block."""
        kg = grids if key_grids is None else key_grids
        key = f"dst_rows:{nx}x{ny}:p{precision}:g{kg}"
        if key in self.choices:
            return self.choices[key]
        m = 16
        while m < 2 * nx - 1:
            m *= 2
        esize = 8 if precision == N.FB_F32 else 16  # bytes of one complex value in LDS
        slots = m + m // 8  # LDS slots of one row: a pad slot after every 8 points
        times = {}
        for ppt in (16, 8):
            nthr = m // ppt
            if nthr > 512:
                continue
            cands = [r for r in (1, 2, 4, 8, 16, 32)
                     if r * nthr <= 512 and r * slots * esize <= 96 * 1024 and r <= ny * grids]
            self._set(N.FB_TUNE_DST_PPT, nx, ny, precision, grids, ppt)
            for r in cands:
                self._set(N.FB_TUNE_DST_ROWS, nx, ny, precision, grids, r)
                ms = C.c_double()
                N.check(self.lib.fb_poisson_time_precond(self.ctx.device, nx, ny, dx, dy, precision, grids, self.reps,
                                                         C.byref(ms)))
                times[f"ppt{ppt}:rows{r}"] = ms.value
        self._set(N.FB_TUNE_DST_ROWS, nx, ny, precision, grids, -1)
        self._set(N.FB_TUNE_DST_PPT, nx, ny, precision, grids, -1)
        best = min(times, key=times.get)
        ppt, rows = int(best.split(":")[0][3:]), int(best.split(":")[1][4:])
        self._set(N.FB_TUNE_DST_ROWS, nx, ny, precision, kg, rows)
        self._set(N.FB_TUNE_DST_PPT, nx, ny, precision, kg, ppt)
        rec = {"param": N.FB_TUNE_DST_ROWS, "nx": nx, "ny": ny, "precision": precision, "grids": kg, "value": rows,
               "ppt": ppt, "ms": times, "set": [[N.FB_TUNE_DST_PPT, nx, ny, precision, kg, ppt]]}
        self.choices[key] = rec
        return rec

    # ------------------------------------------------------------------ lattice step shape
    def lbm(self, nx: int, ny: int, form: str = "f32-mixed", steps: int = 100) -> dict:
        """The single-step shapes (interior rows per thread x nontemporal population traffic) and the fused pair,
        timed jointly (rows and streaming interact: two rows per thread pay with cached loads, one row with
        streaming ones), then the blocks striding over the flagged cells of the winner, for one lattice shape and
        storage form."""
        from .benchmark.lbm_harness import Lattice, LatticeCase

        prec = _PREC[form]
        key = f"lbm:{nx}x{ny}:p{prec}"
        if key in self.choices:
            return self.choices[key]
        lat = Lattice(LatticeCase(nx=nx, ny=ny), prec, self.ctx)

        def timed() -> float:
            lat.time(4)
            return min(lat.time(steps) / steps for _ in range(3))

        times = {}
        try:
            shapes = [(0, rows, nt) for rows in (1, 2) for nt in (0, 1)] + [(1, 1, -1)]
            for fuse, rows, nt in shapes:
                self._set(N.FB_TUNE_LBM_FUSE, nx, ny, prec, 1, fuse)
                self._set(N.FB_TUNE_LBM_ROWS, nx, ny, prec, 1, rows)
                self._set(N.FB_TUNE_LBM_NT, nx, ny, prec, 1, nt)
                times[f"fuse{fuse}:rows{rows}:nt{nt}"] = timed()
            best = min(times, key=times.get)
            fuse, rows, nt = (int(t[len(n):]) for t, n in zip(best.split(":"), ("fuse", "rows", "nt")))
            self._set(N.FB_TUNE_LBM_FUSE, nx, ny, prec, 1, fuse)
            self._set(N.FB_TUNE_LBM_ROWS, nx, ny, prec, 1, rows)
            self._set(N.FB_TUNE_LBM_NT, nx, ny, prec, 1, nt)
            list_blocks = -1
            if not fuse:
                lt = {}
                for b in (8, 16, 32, 64, 128):
                    self._set(N.FB_TUNE_LBM_LIST_BLOCKS, nx, ny, prec, 1, b)
                    lt[b] = timed()
                list_blocks = min(lt, key=lt.get)
                times.update({f"list{b}": v for b, v in lt.items()})
                self._set(N.FB_TUNE_LBM_LIST_BLOCKS, nx, ny, prec, 1, list_blocks)
        finally:
            lat.close()
        rec = {"fuse": fuse, "rows": rows, "list_blocks": list_blocks, "nt": nt, "ms": times,
               "set": [[N.FB_TUNE_LBM_FUSE, nx, ny, prec, 1, fuse], [N.FB_TUNE_LBM_ROWS, nx, ny, prec, 1, rows],
                       [N.FB_TUNE_LBM_LIST_BLOCKS, nx, ny, prec, 1, list_blocks],
                       [N.FB_TUNE_LBM_NT, nx, ny, prec, 1, nt]]}
        self.choices[key] = rec
        return rec

    # ------------------------------------------------------------------ lattice batch chunk
    def lbm_chunk(self, shapes: list[tuple[int, int]], form: str = "f32-mixed", steps: int = 64) -> dict:
        from .benchmark.lbm_harness import Lattice, LatticeCase

        key = "lbm_chunk:" + ",".join(f"{nx}x{ny}" for nx, ny in shapes) + f":{form}"
        if key in self.choices:
            return self.choices[key]
        prec = _PREC[form]
        lats = [Lattice(LatticeCase(nx=nx, ny=ny), prec, self.ctx) for nx, ny in shapes]
        times = {}
        try:
            handles = (C.c_void_p * len(lats))(*[l.handle for l in lats])
            params = (N.fb_lbm_params * len(lats))(*[l.params for l in lats])
            for chunk in (8, 16, 32, 64):
                N.check(self.lib.fb_lbm_run_batch(handles, params, len(lats), chunk, chunk, None))
                t0 = time.perf_counter()
                N.check(self.lib.fb_lbm_run_batch(handles, params, len(lats), steps, chunk, None))
                times[chunk] = (time.perf_counter() - t0) * 1e3 / steps
        finally:
            for l in lats:
                l.close()
        best = min(times, key=times.get)
        rec = {"value": best, "ms_per_step": times}
        self.choices[key] = rec
        return rec

    # ------------------------------------------------------------------ a request mix
    def for_mix(self, requests, sf_batch: int) -> dict:
        """Tunes every grid shape of a serving mix; returns the request loop's lattice chunk."""
        from .serving import engine
        from .serving.mix import LBM, SF

        # the precision of the preconditioner the serving form's Stable Fluids requests apply
        sf_prec = N.FB_F32 if engine.SF_SOLVER_FORM == "precond32" else N.FB_F64

        sf_shapes: dict[tuple[int, int], int] = {}
        lbm_shapes: set[tuple[int, int]] = set()
        for r in requests:
            if r.solver == SF:
                sf_shapes[(r.nx, r.ny)] = sf_shapes.get((r.nx, r.ny), 0) + 1
            elif r.solver == LBM:
                lbm_shapes.add((r.nx, r.ny))
        out = {"dst_rows": {}, "lbm": {}}
        for (nx, ny), count in sorted(sf_shapes.items()):
            dx = 4.0 / nx
            g = max(1, min(sf_batch, count, N_MAX_BATCH))
            out["dst_rows"][f"{nx}x{ny}"] = self.poisson(nx, ny, dx, dx, sf_prec, grids=g, key_grids=0)["value"]
        for nx, ny in sorted(lbm_shapes):
            r = self.lbm(nx, ny)
            out["lbm"][f"{nx}x{ny}"] = {k: r[k] for k in ("fuse", "rows", "list_blocks")}
        out["lbm_chunk"] = self.lbm_chunk(sorted(lbm_shapes))["value"] if lbm_shapes else 32
        self.save()
        return out


N_MAX_BATCH = 64
