"""Blended PIC/FLIP restated from kern/pic.py, fb/solvers/_numba_adapter.py:108-160,
fb/solvers/pic_flip.py and fb/core/particle_population.py (test infrastructure only).

Transfers (face-centred, and the cell-centred pair of kern/pic.py:22-102) are vectorised but
keep the reference's accumulation order (particle-major, then the 3x3 offsets, via unbuffered
np.add.at into accumulators of the positions' dtype) and its FP64 weight arithmetic.

Pinned by (tests/test_oracle_pic.py): a per-particle loop restating the numba kernels on
tiny inputs, constant preservation and periodic wrap of the transfers
(tests/unit/test_pic_transfers.py:80-258), partition of unity of the quadratic B-spline,
the population-maintenance cases of tests/unit/test_particle_population.py, and the PCG32
fixture for seeding draws.
"""

from __future__ import annotations

import numpy as np

from .geometry import NacaFoil, cell_centers
from .lbm import OracleFailure
from .mac import Dom, cell_to_faces, faces_to_cell, implicit_diffuse_scalar, native_divergence_linf, \
    project_faces, solid_face_leakage
from .rng import PCG32


def bspline(d):
    """kern/pic.py:11-19 (vectorised)."""
    a = np.abs(d)
    return np.where(a < 0.5, 0.75 - a * a, np.where(a < 1.5, 0.5 * (1.5 - a) * (1.5 - a), 0.0))


def _support(positions, x0, y0, dx, dy, gxo, gyo):
    gx = (positions[:, 0].astype(np.float64) - x0) / dx + gxo
    gy = (positions[:, 1].astype(np.float64) - y0) / dy + gyo
    bx = np.floor(gx - 0.5).astype(np.int64)
    by = np.floor(gy - 0.5).astype(np.int64)
    return gx, gy, bx, by


def scatter_face_component(positions, values, x0, y0, dx, dy, width, height, gxo, gyo, px, py, dupx, dupy,
                           fallback):
    """kern/pic.py:105-161."""
    dt = positions.dtype
    w = np.zeros(height * width, dtype=dt)
    m = np.zeros(height * width, dtype=dt)
    uw = width - 1 if dupx else width
    uh = height - 1 if dupy else height
    gx, gy, bx, by = _support(positions, x0, y0, dx, dy, gxo, gyo)
    idx = []
    wt = []
    for oy in range(3):
        sy = by + oy
        ty = np.mod(sy, uh) if py else sy
        wy = bspline(gy - sy)
        for ox in range(3):
            sx = bx + ox
            tx = np.mod(sx, uw) if px else sx
            ok = (ty >= 0) & (ty < height) & (tx >= 0) & (tx < width)
            idx.append(np.where(ok, ty * width + tx, -1))
            wt.append(wy * bspline(gx - sx))
    idx = np.stack(idx, axis=1).ravel()   # particle-major, then oy, ox (the numba loop order)
    wt = np.stack(wt, axis=1).ravel()
    vals = np.repeat(values.astype(np.float64), 9)
    keep = idx >= 0
    np.add.at(w, idx[keep], wt[keep])
    np.add.at(m, idx[keep], wt[keep] * vals[keep])
    w = w.reshape(height, width)
    m = m.reshape(height, width)
    out = np.where(w > 1.0e-12, m / np.where(w > 1.0e-12, w, 1), fallback).astype(dt)
    unsupported = int(np.count_nonzero(~(w > 1.0e-12)))
    if dupx:
        out[:, width - 1] = out[:, 0]
    if dupy:
        out[height - 1, :] = out[0, :]
    return out, unsupported


def gather_face_component(field, positions, x0, y0, dx, dy, gxo, gyo, px, py, dupx, dupy):
    """kern/pic.py:164-205 (FP64 accumulation in the fixed 3x3 order)."""
    height, width = field.shape
    uw = width - 1 if dupx else width
    uh = height - 1 if dupy else height
    gx, gy, bx, by = _support(positions, x0, y0, dx, dy, gxo, gyo)
    val = np.zeros(positions.shape[0])
    ws = np.zeros(positions.shape[0])
    f = field.astype(np.float64)
    for oy in range(3):
        sy = by + oy
        ty = np.mod(sy, uh) if py else sy
        wy = bspline(gy - sy)
        for ox in range(3):
            sx = bx + ox
            tx = np.mod(sx, uw) if px else sx
            ok = (ty >= 0) & (ty < height) & (tx >= 0) & (tx < width)
            w = wy * bspline(gx - sx)
            v = np.where(ok, f[np.clip(ty, 0, height - 1), np.clip(tx, 0, width - 1)], 0.0)
            val = np.where(ok, val + w * v, val)
            ws = np.where(ok, ws + w, ws)
    return (val / np.maximum(ws, 1.0e-12)).astype(positions.dtype)


def particle_to_faces(positions, velocities, fu, fv, x0, y0, dx, dy, px=False, py=False):
    """fb/solvers/_numba_adapter.py:108-135."""
    u, nu = scatter_face_component(positions, velocities[:, 0], x0, y0, dx, dy, fu.shape[1], fu.shape[0], 0.0, -0.5,
                                   px, py, px, False, fu)
    v, nv = scatter_face_component(positions, velocities[:, 1], x0, y0, dx, dy, fv.shape[1], fv.shape[0], -0.5, 0.0,
                                   px, py, False, py, fv)
    return u, v, (nu + nv) / (u.size + v.size)


def faces_to_particle(u, v, positions, x0, y0, dx, dy, px=False, py=False):
    """fb/solvers/_numba_adapter.py:138-160."""
    out = np.empty_like(positions)
    out[:, 0] = gather_face_component(u, positions, x0, y0, dx, dy, 0.0, -0.5, px, py, px, False)
    out[:, 1] = gather_face_component(v, positions, x0, y0, dx, dy, -0.5, 0.0, px, py, False, py)
    return out



def _cell_stencil(positions, x0, y0, dx, dy, nx, ny, px, py):
    """Targets and FP64 weights of the cell-centred 3x3 stencil, particle-major then (oy, ox) (kern/pic.py:40-53,
    84-94): the stencil index wrapped on a periodic axis, clamped to [0, n - 1] otherwise."""
    gx, gy, bx, by = _support(positions, x0, y0, dx, dy, -0.5, -0.5)
    tgt, wt = [], []
    for oy in range(3):
        sy = by + oy
        ty = np.mod(sy, ny) if py else np.clip(sy, 0, ny - 1)
        wy = bspline(gy - sy)
        for ox in range(3):
            sx = bx + ox
            tx = np.mod(sx, nx) if px else np.clip(sx, 0, nx - 1)
            tgt.append(ty * nx + tx)
            wt.append(wy * bspline(gx - sx))
    return np.stack(tgt, axis=1), np.stack(wt, axis=1)


def particle_to_grid(positions, velocities, x0, y0, dx, dy, nx, ny, freestream, px=False, py=False):
    """kern/pic.py:22-65 behind fb/solvers/_numba_adapter.py:35-71: weights and momenta accumulated in the
    positions' dtype in particle order (each FP64 contribution added in FP64 and rounded back), the cell velocity
    their dtype quotient where the weight exceeds 1e-12 and the freestream elsewhere."""
    dt = positions.dtype
    tgt, wt = _cell_stencil(positions, x0, y0, dx, dy, nx, ny, px, py)
    idx = tgt.ravel()
    wt = wt.ravel()
    w = np.zeros(ny * nx, dtype=dt)
    mx = np.zeros(ny * nx, dtype=dt)
    my = np.zeros(ny * nx, dtype=dt)
    np.add.at(w, idx, wt)
    np.add.at(mx, idx, wt * np.repeat(velocities[:, 0].astype(np.float64), 9))
    np.add.at(my, idx, wt * np.repeat(velocities[:, 1].astype(np.float64), 9))
    ok = w > 1.0e-12
    den = np.where(ok, w, dt.type(1))
    grid = np.empty((ny, nx, 2), dtype=dt)
    grid[..., 0] = np.where(ok, mx / den, dt.type(freestream[0])).reshape(ny, nx)
    grid[..., 1] = np.where(ok, my / den, dt.type(freestream[1])).reshape(ny, nx)
    return grid


def grid_to_particle(grid, positions, x0, y0, dx, dy, px=False, py=False):
    """kern/pic.py:67-102 behind fb/solvers/_numba_adapter.py:74-105: FP64 sums in the fixed 3x3 order, times
    1 / max(weight sum, 1e-12), rounded to the dtype."""
    ny, nx, _ = grid.shape
    tgt, wt = _cell_stencil(positions, x0, y0, dx, dy, nx, ny, px, py)
    g = grid.reshape(ny * nx, 2).astype(np.float64)
    vx = np.zeros(positions.shape[0])
    vy = np.zeros(positions.shape[0])
    ws = np.zeros(positions.shape[0])
    for k in range(9):
        vx = vx + wt[:, k] * g[tgt[:, k], 0]
        vy = vy + wt[:, k] * g[tgt[:, k], 1]
        ws = ws + wt[:, k]
    inv = 1.0 / np.maximum(ws, 1.0e-12)
    return np.stack((vx * inv, vy * inv), axis=1).astype(positions.dtype)

# ---------------------------------------------------------------- population maintenance

def particle_cell_ids(positions, dom):
    """fb/core/particle_population.py:19-31."""
    cx = np.floor((positions[:, 0] - dom.bounds[0][0]) / dom.dx).astype(np.int64)
    cy = np.floor((positions[:, 1] - dom.bounds[1][0]) / dom.dy).astype(np.int64)
    cx = np.clip(cx, 0, dom.nx - 1)
    cy = np.clip(cy, 0, dom.ny - 1)
    return cy * dom.nx + cx


def particle_cell_counts(positions, dom):
    return np.bincount(particle_cell_ids(positions, dom), minlength=dom.nx * dom.ny).astype(np.int64)


def maintain_particle_population(positions, solid, dom, rng, minimum=2, target=4, maximum=8):
    """fb/core/particle_population.py:44-135. Returns (moved indices, counts)."""
    n = dom.nx * dom.ny
    ids = particle_cell_ids(positions, dom)
    counts = np.bincount(ids, minlength=n).astype(np.int64)
    fluid_ids = np.flatnonzero(~solid.reshape(-1))
    if fluid_ids.size == 0:
        raise ValueError("particle population requires at least one fluid cell")
    fc = counts[fluid_ids]
    solid_occ = bool(np.any(counts[solid.reshape(-1)]))
    if bool(np.all((fc >= minimum) & (fc <= maximum))) and not solid_occ:
        return np.empty(0, dtype=np.int64), counts
    desired = counts.copy()
    desired[solid.reshape(-1)] = 0
    desired[fluid_ids[counts[fluid_ids] < minimum]] = target
    desired[fluid_ids[counts[fluid_ids] > maximum]] = target
    diff = positions.shape[0] - int(np.sum(desired))
    if diff > 0:
        cap = maximum - desired[fluid_ids]
        pri = np.lexsort((fluid_ids, desired[fluid_ids]))
        slots = np.repeat(fluid_ids[pri], cap[pri])
        if slots.size < diff:
            raise RuntimeError("particle count exceeds configured maximum population")
        np.add.at(desired, slots[:diff], 1)
    elif diff < 0:
        rem = desired[fluid_ids] - minimum
        pri = np.lexsort((fluid_ids, -desired[fluid_ids]))
        slots = np.repeat(fluid_ids[pri], rem[pri])
        if slots.size < -diff:
            raise RuntimeError("particle count cannot satisfy configured minimum population")
        np.subtract.at(desired, slots[:-diff], 1)
    order = np.argsort(ids, kind="stable")
    sc = ids[order]
    seq = np.arange(order.size, dtype=np.int64)
    gs = np.zeros(order.size, dtype=np.int64)
    new = np.empty(order.size, dtype=bool)
    new[0] = True
    new[1:] = sc[1:] != sc[:-1]
    gs[new] = seq[new]
    np.maximum.accumulate(gs, out=gs)
    rank = seq - gs
    donors = order[rank >= desired[sc]]
    deficits = desired - np.minimum(counts, desired)
    dest = np.repeat(np.arange(n, dtype=np.int64), deficits)
    if donors.size != dest.size:
        raise RuntimeError("particle population redistribution is unbalanced")
    if donors.size:
        dxc = dest % dom.nx
        dyc = dest // dom.nx
        jit = rng.random((donors.size, 2)).astype(positions.dtype, copy=False)
        positions[donors, 0] = dom.bounds[0][0] + (dxc + 0.4 + 0.2 * jit[:, 0]) * dom.dx
        positions[donors, 1] = dom.bounds[1][0] + (dyc + 0.4 + 0.2 * jit[:, 1]) * dom.dy
    return donors, particle_cell_counts(positions, dom)


# ---------------------------------------------------------------- solver

def seed_particles(solid, dom, rng, dtype, per_cell=4):
    """fb/solvers/pic_flip.py:317-333 (per_cell is 4 in the reference)."""
    fy, fx = np.nonzero(~solid)
    cx = np.repeat(fx, per_cell)
    cy = np.repeat(fy, per_cell)
    jit = rng.random((cx.size, 2)).astype(dtype)
    pos = np.empty((cx.size, 2), dtype=dtype)
    pos[:, 0] = dom.bounds[0][0] + (cx + 0.1 + 0.8 * jit[:, 0]) * dom.dx
    pos[:, 1] = dom.bounds[1][0] + (cy + 0.1 + 0.8 * jit[:, 1]) * dom.dy
    return pos


class PicFlipOracle:
    """Restates ``PicFlipSolver`` (fb/solvers/pic_flip.py:135-1008)."""

    def initialize(self, sc, seed=0, per_cell=4):
        # fb/solvers/pic_flip.py:248-288
        self.sc = sc
        self.dom = Dom.of(sc.domain)
        self.foil = NacaFoil(sc.foil.naca, sc.foil.chord, tuple(sc.foil.pivot[:2]))
        self.control = sc.control_at(0.0)
        self.time = 0.0
        self.reynolds = float(sc.reynolds)
        self.blend = float(np.clip(float(sc.solver_options.get("pic_flip_blend", 0.95)), 0.0, 1.0))
        self.interval = int(sc.solver_options.get("pic_population_interval", 8))
        self.cfl = float(sc.solver_options.get("pic_cfl", 0.75))
        self.advance_count = 0
        self.settling = 0
        self.rng = PCG32(seed, stream=71)
        self.solid = self.foil.mask(self.dom.bounds, self.dom.nx, self.dom.ny, self.control.angle_degrees)
        self.solid_angle = self.control.angle_degrees
        from .mac import initial_velocity
        vel = initial_velocity(sc)
        self.grid = vel
        self.per_cell = per_cell
        self.positions = seed_particles(self.solid, self.dom, self.rng, sc.dtype, per_cell)
        u, v = cell_to_faces(vel)
        self.velocities = self._f2p(u, v, self.positions)
        # fb/solvers/pic_flip.py:287: the P2G result is discarded; only the unsupported fraction is kept
        self._p2g_cell()

    def import_state(self, velocity, time, control):
        # fb/solvers/pic_flip.py:1035-1113: the grid velocity with the solid cells at the wall velocity of the
        # import pose, particles reseeded from the stream where it stands (fb/solvers/pic_flip.py:317-333) with
        # velocities gathered from that grid, and one settling advance (blend 0) to follow
        saved = self._checkpoint()
        try:
            vel = np.asarray(velocity, dtype=self.sc.dtype).copy()
            self.time = float(time)
            self.control = control
            self.solid = self.foil.mask(self.dom.bounds, self.dom.nx, self.dom.ny, control.angle_degrees)
            self.solid_angle = control.angle_degrees
            wall = self._wall(control.angular_velocity_degrees)
            vel[self.solid] = wall[self.solid]
            self.grid = vel
            self.positions = seed_particles(self.solid, self.dom, self.rng, self.sc.dtype, self.per_cell)
            u, v = cell_to_faces(self.grid)
            self.velocities = self._f2p(u, v, self.positions)
            self.settling = 1
        except Exception:
            self._restore(saved)
            raise

    def _px(self):
        return "x" in self.dom.periodic_axes, "y" in self.dom.periodic_axes

    def _f2p(self, u, v, pos):
        px, py = self._px()
        return faces_to_particle(u, v, pos, self.dom.bounds[0][0], self.dom.bounds[1][0], self.dom.dx, self.dom.dy,
                                 px, py)

    def _p2f(self, fu, fv):
        px, py = self._px()
        u, v, uns = particle_to_faces(self.positions, self.velocities, fu, fv, self.dom.bounds[0][0],
                                      self.dom.bounds[1][0], self.dom.dx, self.dom.dy, px, py)
        self.unsupported = uns
        return u, v

    def _p2g_cell(self):
        fu, fv = cell_to_faces(self.grid)
        u, v = self._p2f(fu, fv)
        return faces_to_cell(u, v)

    def _wall(self, angular_velocity_degrees):
        pts = cell_centers(self.dom.bounds, self.dom.nx, self.dom.ny).reshape(-1, 2)
        return self.foil.wall_velocity(pts, angular_velocity_degrees).reshape(self.dom.ny, self.dom.nx, 2)

    def resolve_collisions(self, control):
        # fb/solvers/pic_flip.py:604-618
        pos = self.positions
        inside = self.foil.contains(pos, control.angle_degrees)
        if np.any(inside):
            pts = pos[inside]
            d = self.foil.signed_distance(pts, control.angle_degrees)
            n = self.foil.normals(pts, control.angle_degrees)
            pos[inside] -= (d[:, None] - 1.0e-4) * n
            wall = self.foil.wall_velocity(pos[inside], control.angular_velocity_degrees)
            rel = self.velocities[inside] - wall
            into = np.sum(rel * n, axis=1) < 0.0
            rel[into] -= np.sum(rel[into] * n[into], axis=1)[:, None] * n[into]
            self.velocities[inside] = wall + rel

    def swept_collisions(self, start, start_control, control):
        # fb/solvers/pic_flip.py:527-602
        pos = self.positions
        path = pos - start
        pivot = np.asarray(self.foil.pivot, dtype=pos.dtype)
        l2 = np.sum(path * path, axis=1)
        proj = np.sum((pivot - start) * path, axis=1) / np.maximum(l2, 1.0e-12)
        proj = np.clip(proj, 0.0, 1.0)
        closest = start + proj[:, None] * path
        margin = 0.05 * min(self.dom.dx, self.dom.dy)
        cand = np.linalg.norm(closest - pivot, axis=1) <= (self.foil.maximum_radius + margin)
        ci = np.flatnonzero(cand)
        if ci.size == 0:
            return 0
        travel = float(np.max(np.linalg.norm(path[ci], axis=1)))
        wall_travel = abs(np.deg2rad(control.angle_degrees - start_control.angle_degrees)) * self.foil.maximum_radius
        spacing = 0.1 * min(self.dom.dx, self.dom.dy)
        count = int(np.clip(np.ceil((travel + wall_travel) / max(spacing, 1.0e-12)), 2, 16))
        hit = np.zeros(ci.size, dtype=bool)
        for s in range(1, count + 1):
            frac = s / count
            angle = start_control.angle_degrees + frac * (control.angle_degrees - start_control.angle_degrees)
            sp = start[ci] + frac * path[ci]
            d = self.foil.signed_distance(sp, angle)
            ent = (d <= margin) & ~hit
            if not np.any(ent):
                continue
            sel = ci[ent]
            ip = sp[ent]
            idist = d[ent]
            n = self.foil.normals(ip, angle)
            res = ip - (idist[:, None] - margin) * n
            pos[sel] = res
            wall = self.foil.wall_velocity(res, control.angular_velocity_degrees)
            rel = self.velocities[sel] - wall
            inward = np.sum(rel * n, axis=1)
            into = inward < 0.0
            rel[into] -= inward[into, None] * n[into]
            self.velocities[sel] = wall + rel
            hit[ent] = True
        return int(np.count_nonzero(hit))

    def advect(self, start_control, control, dt, v0, u, v):
        # fb/solvers/pic_flip.py:481-525
        pos = self.positions
        start = pos.copy()
        mid = pos + 0.5 * dt * v0
        vm = self._f2p(u, v, mid)
        pos += dt * vm
        self.swept_collisions(start, start_control, control)
        (x0, x1), (y0, y1) = self.dom.bounds
        px, py = self._px()
        if px:
            pos[:, 0] = x0 + np.mod(pos[:, 0] - x0, x1 - x0)
        if py:
            pos[:, 1] = y0 + np.mod(pos[:, 1] - y0, y1 - y0)
        esc = np.zeros(pos.shape[0], dtype=bool)
        if not px:
            esc |= (pos[:, 0] < x0) | (pos[:, 0] > x1)
        if not py:
            esc |= (pos[:, 1] < y0) | (pos[:, 1] > y1)
        count = int(np.count_nonzero(esc))
        if count:
            if px:
                pos[esc, 0] = self.rng.uniform(x0, x1, (count,))
            else:
                pos[esc, 0] = x0 + self.rng.uniform(0.0, 0.5 * self.dom.dx, (count,))
            pos[esc, 1] = self.rng.uniform(y0, y1, (count,))
            self.velocities[esc] = np.asarray(self.sc.freestream[:2], dtype=self.sc.dtype)
        self.resolve_collisions(control)

    def _checkpoint(self):
        # fb/solvers/pic_flip.py:737-755
        return (self.positions.copy(), self.velocities.copy(), self.grid.copy(), self.solid.copy(), self.solid_angle,
                self.control, self.time, (self.rng.state, self.rng.inc), self.advance_count, self.settling,
                getattr(self, "unsupported", 0.0))

    def _restore(self, c):
        (self.positions, self.velocities, self.grid, self.solid, self.solid_angle, self.control, self.time, words,
         self.advance_count, self.settling, self.unsupported) = c
        self.rng.state, self.rng.inc = words

    def advance(self, control, target_dt):
        # fb/solvers/pic_flip.py:637-677: retry with more substeps when the post-step particle CFL exceeds pic_cfl
        minimum = 1
        retries = 0
        while True:
            try:
                return self.advance_once(control, target_dt, minimum, retries)
            except OracleFailure as err:
                acc = err.evidence.get("accepted_cfl")
                mx = err.evidence.get("maximum_cfl")
                att = err.evidence.get("substeps")
                if not (err.reason == "stability_limit" and isinstance(acc, float) and isinstance(mx, float)
                        and isinstance(att, int) and acc > mx > 0.0):
                    raise
                nxt = max(att + 1, int(np.ceil(1.05 * att * acc / mx)))
                if nxt > 512:
                    raise OracleFailure("stability_limit", "retry requires too many substeps", "particle-advection",
                                        {**err.evidence, "required_substeps": nxt, "stability_retries": retries})
                retries += 1
                minimum = nxt

    def advance_once(self, control, target_dt, minimum_substeps=1, retries=0):
        # fb/solvers/pic_flip.py:679-1008
        saved = self._checkpoint()
        try:
            return self._advance_body(control, target_dt, minimum_substeps, retries)
        except Exception:
            self._restore(saved)
            raise

    def _advance_body(self, control, target_dt, minimum_substeps, retries):
        sc = self.sc
        dom = self.dom
        max_speed = max(float(np.max(np.linalg.norm(self.grid, axis=2))),
                        float(np.max(np.linalg.norm(self.velocities, axis=1))), abs(sc.freestream[0]), 1.0e-6)
        bav = control.angular_velocity_degrees
        sweep_av = (control.angle_degrees - self.control.angle_degrees) / target_dt
        wall_speed = self.foil.maximum_radius * max(abs(np.deg2rad(bav)), abs(np.deg2rad(sweep_av)))
        transport = max(max_speed, wall_speed)
        stable_dt = self.cfl * min(dom.dx, dom.dy) / max(transport, 1.0e-6)
        subs = max(minimum_substeps, int(np.ceil(target_dt / stable_dt)))
        if subs > 512:
            raise OracleFailure("stability_limit", "PIC/FLIP motion requires too many internal substeps",
                                "particle-advection", {"required_substeps": subs})
        dt = target_dt / subs
        start_time = self.time
        start_angle = self.control.angle_degrees
        nu = sc.reference_speed * sc.foil.chord / self.reynolds
        tol = float(sc.solver_options.get("pressure_tolerance", 1.0e-5))
        maxit = int(sc.solver_options.get("pressure_max_iterations", 640))
        channel = str(sc.solver_options.get("initial_condition", "")) == "poiseuille"
        for k in range(subs):
            frac = (k + 1) / subs
            sub = type(control)(start_time + frac * target_dt, start_angle + frac * (control.angle_degrees - start_angle),
                                bav)
            if self.solid_angle != sub.angle_degrees:
                self.solid = self.foil.mask(dom.bounds, dom.nx, dom.ny, sub.angle_degrees)
                self.solid_angle = sub.angle_degrees
            start_control = self.control
            self.resolve_collisions(sub)
            fu, fv = cell_to_faces(self.grid)
            tu, tv = self._p2f(fu, fv)
            pu, pv = tu.copy(), tv.copy()
            du, its_u, ru, oku = implicit_diffuse_scalar(tu, nu, dt, dom, tol, maxit)
            dv, its_v, rv, okv = implicit_diffuse_scalar(tv, nu, dt, dom, tol, maxit)
            self.last_viscosity = (max(its_u, its_v), max(ru, rv), oku and okv)
            wall = self._wall(sub.angular_velocity_degrees)
            qu, qv, rep, _ = project_faces(du, dv, dom, self.solid, wall, sc.freestream, dt, channel, tol, maxit)
            self.last_projection = rep
            self.grid = faces_to_cell(qu, qv)
            pic = self._f2p(qu, qv, self.positions)
            delta = self._f2p(qu - pu, qv - pv, self.positions)
            b = 0.0 if self.settling > 0 else self.blend
            self.velocities[:] = (1.0 - b) * pic + b * (self.velocities + delta)
            self.control = sub
            self.advect(start_control, sub, dt, pic, qu, qv)
        if self.settling > 0:
            self.settling -= 1
        self.advance_count += 1
        if self.advance_count % self.interval == 0:
            moved, _ = maintain_particle_population(self.positions, self.solid, dom, self.rng)
            if moved.size:
                fu, fv = cell_to_faces(self.grid)
                self.velocities[moved] = self._f2p(fu, fv, self.positions[moved])
                self.resolve_collisions(control)
        self.resolve_collisions(control)
        if not (np.isfinite(self.grid).all() and np.isfinite(self.positions).all()
                and np.isfinite(self.velocities).all()):
            raise OracleFailure("nonfinite_state", "PIC/FLIP produced non-finite state", "postcondition")
        inside = self.foil.contains(self.positions, control.angle_degrees)
        if np.any(inside):
            raise OracleFailure("postcondition_failure", "particles inside the moving solid", "postcondition",
                                {"unresolved_solid_particles": int(np.count_nonzero(inside))})
        # fb/solvers/pic_flip.py:872-893: the accepted particle CFL of the step
        report_speed = max(transport, float(np.max(np.linalg.norm(self.grid, axis=2))),
                           float(np.max(np.linalg.norm(self.velocities, axis=1))))
        accepted = dt * report_speed / min(dom.dx, dom.dy)
        if accepted > self.cfl * (1.0 + 1.0e-6):
            raise OracleFailure("stability_limit", "post-step motion exceeded its swept envelope", "particle-advection",
                                {"accepted_cfl": accepted, "maximum_cfl": self.cfl, "substeps": subs})
        self.time = start_time + target_dt
        self.control = type(control)(self.time, control.angle_degrees, control.angular_velocity_degrees)
        wallf = self._wall(control.angular_velocity_degrees)
        return {"substeps": subs, "pressure_iterations": self.last_projection[0],
                "pressure_relative_residual": self.last_projection[1],
                "viscosity_iterations": self.last_viscosity[0],
                "viscosity_final_residual": self.last_viscosity[1],
                "divergence_linf": native_divergence_linf(qu, qv, dom, self.solid),
                "solid_leakage": solid_face_leakage(qu, qv, self.solid, wallf),
                "unsupported_face_fraction": self.unsupported, "stability_retries": retries,
                "maximum_particle_cfl": accepted, "particle_count": int(self.positions.shape[0])}
