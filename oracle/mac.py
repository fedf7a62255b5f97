"""Staggered MAC-grid operations and the Stable Fluids solver, restated from fb/core/grid.py,
fb/core/interpolation.py, fb/core/_scipy_adapter.py and fb/solvers/stable_fluids.py.

Test infrastructure only. The pressure solve uses SciPy's own ``cg`` and ``dstn``/``idstn``
(the third-party routines the reference calls; the reference pins SciPy 1.18.0, this
container has 1.15.3 whose ``cg`` loop is the same: convergence test ||r|| < max(atol,
rtol*||b||) at the top of each iteration, fb/core/_scipy_adapter.py:132-147).

Pinned by (tests/test_oracle_mac.py): the MAC boundary fixture (spec/conformance/
mac-boundary.json and tests/conformance/test_revision5_conformance.py:169-185), a dense
direct solve of the masked operator on tiny grids, exactness of the DST preconditioner on
a solid-free box (CG converges in one iteration), the discrete Fourier decay of implicit
diffusion (tests/unit/test_projection.py:90-112), the RK2 rigid-rotation closed form
(:146-160), constant preservation and the MacCormack bound property (:163-208).
"""

from __future__ import annotations

import numpy as np
from scipy.fft import dstn, idstn
from scipy.sparse.linalg import LinearOperator, cg

from .geometry import NacaFoil, cell_centers
from .lbm import OracleFailure


class Dom:
    """Minimal domain description (nx, ny, bounds, periodic axes)."""

    def __init__(self, bounds, nx, ny, periodic_axes=()):
        self.bounds = tuple(tuple(float(v) for v in b) for b in bounds)
        self.nx = int(nx)
        self.ny = int(ny)
        self.periodic_axes = tuple(periodic_axes)
        self.dx = (self.bounds[0][1] - self.bounds[0][0]) / self.nx
        self.dy = (self.bounds[1][1] - self.bounds[1][0]) / self.ny

    @classmethod
    def of(cls, d):
        return cls(d.bounds, d.nx, d.ny, d.periodic_axes)


# ----------------------------------------------------------------------------- faces

def faces_to_cell(u, v):
    """fb/core/grid.py:25-29."""
    return np.stack((0.5 * (u[:, :-1] + u[:, 1:]), 0.5 * (v[:-1, :] + v[1:, :])), axis=-1)


def cell_to_faces(vel):
    """fb/core/grid.py:32-44."""
    ny, nx, _ = vel.shape
    u = np.empty((ny, nx + 1), dtype=vel.dtype)
    v = np.empty((ny + 1, nx), dtype=vel.dtype)
    u[:, 1:-1] = 0.5 * (vel[:, :-1, 0] + vel[:, 1:, 0])
    u[:, 0] = vel[:, 0, 0]
    u[:, -1] = vel[:, -1, 0]
    v[1:-1, :] = 0.5 * (vel[:-1, :, 1] + vel[1:, :, 1])
    v[0, :] = vel[0, :, 1]
    v[-1, :] = vel[-1, :, 1]
    return u, v


def apply_domain_boundaries(u, v, dom, freestream, channel_walls=False):
    """fb/core/grid.py:47-78."""
    ux, uy = freestream[0], freestream[1]
    if "x" in dom.periodic_axes:
        pu = 0.5 * (u[:, 0] + u[:, -1])
        u[:, 0] = pu
        u[:, -1] = pu
    else:
        u[:, 0] = ux
        u[:, -1] = u[:, -2]
        v[:, 0] = uy
        v[:, -1] = v[:, -2]
    if "y" in dom.periodic_axes:
        pv = 0.5 * (v[0, :] + v[-1, :])
        v[0, :] = pv
        v[-1, :] = pv
    elif channel_walls:
        v[0, :] = 0.0
        v[-1, :] = 0.0
        u[0, :] = 0.0
        u[-1, :] = 0.0
    else:
        v[0, :] = uy
        v[-1, :] = uy
        u[0, :] = ux
        u[-1, :] = ux


def solid_faces(solid):
    su = np.zeros((solid.shape[0], solid.shape[1] + 1), dtype=bool)
    sv = np.zeros((solid.shape[0] + 1, solid.shape[1]), dtype=bool)
    su[:, 1:-1] = solid[:, :-1] | solid[:, 1:]
    su[:, 0] = solid[:, 0]
    su[:, -1] = solid[:, -1]
    sv[1:-1, :] = solid[:-1, :] | solid[1:, :]
    sv[0, :] = solid[0, :]
    sv[-1, :] = solid[-1, :]
    return su, sv


def enforce_solid_faces(u, v, solid, wall):
    """fb/core/grid.py:81-97."""
    su, sv = solid_faces(solid)
    wu, wv = cell_to_faces(wall)
    u[su] = wu[su]
    v[sv] = wv[sv]


def native_divergence_linf(u, v, dom, solid):
    """fb/core/grid.py:100-111."""
    div = (u[:, 1:] - u[:, :-1]) / dom.dx + (v[1:, :] - v[:-1, :]) / dom.dy
    fl = ~solid
    return float(np.max(np.abs(div[fl]))) if np.any(fl) else 0.0


def solid_face_leakage(u, v, solid, wall):
    """fb/core/grid.py:114-131."""
    wu, wv = cell_to_faces(wall)
    iu = np.zeros_like(u, dtype=bool)
    iv = np.zeros_like(v, dtype=bool)
    iu[:, 1:-1] = solid[:, :-1] ^ solid[:, 1:]
    iv[1:-1, :] = solid[:-1, :] ^ solid[1:, :]
    m = 0.0
    if np.any(iu):
        m = max(m, float(np.max(np.abs(u[iu] - wu[iu]))))
    if np.any(iv):
        m = max(m, float(np.max(np.abs(v[iv] - wv[iv]))))
    return m


# ----------------------------------------------------------------------------- Poisson

def masked_operator(fluid, dx, dy, periodic_axes=()):
    """Diagonal and matvec of fb/core/_scipy_adapter.py:27-70."""
    ny, nx = fluid.shape
    ix2 = 1.0 / (dx * dx)
    iy2 = 1.0 / (dy * dy)
    diag = np.zeros((ny, nx))
    he = fluid[:, :-1] & fluid[:, 1:]
    ve = fluid[:-1, :] & fluid[1:, :]
    wx = np.zeros(ny, dtype=bool)
    wy = np.zeros(nx, dtype=bool)
    diag[:, :-1] += he * ix2
    diag[:, 1:] += he * ix2
    diag[:-1, :] += ve * iy2
    diag[1:, :] += ve * iy2
    if "x" in periodic_axes:
        wx = fluid[:, -1] & fluid[:, 0]
        diag[:, -1] += wx * ix2
        diag[:, 0] += wx * ix2
    else:
        diag[:, 0] += fluid[:, 0] * ix2
        diag[:, -1] += fluid[:, -1] * ix2
    if "y" in periodic_axes:
        wy = fluid[-1, :] & fluid[0, :]
        diag[-1, :] += wy * iy2
        diag[0, :] += wy * iy2
    else:
        diag[0, :] += fluid[0, :] * iy2
        diag[-1, :] += fluid[-1, :] * iy2
    diag[~fluid] = 1.0

    def apply(flat):
        p = flat.reshape(ny, nx)
        r = diag * p
        r[:, :-1] -= p[:, 1:] * he * ix2
        r[:, 1:] -= p[:, :-1] * he * ix2
        r[:-1, :] -= p[1:, :] * ve * iy2
        r[1:, :] -= p[:-1, :] * ve * iy2
        if "x" in periodic_axes:
            r[:, -1] -= p[:, 0] * wx * ix2
            r[:, 0] -= p[:, -1] * wx * ix2
        if "y" in periodic_axes:
            r[-1, :] -= p[0, :] * wy * iy2
            r[0, :] -= p[-1, :] * wy * iy2
        r[~fluid] = p[~fluid]
        return r.ravel()

    return diag, apply


def dst_preconditioner(fluid, dx, dy):
    """DST-I fast-Poisson preconditioner (fb/core/_scipy_adapter.py:76-96)."""
    ny, nx = fluid.shape
    ex = 2.0 * (1.0 - np.cos(np.pi * np.arange(1, nx + 1, dtype=np.float64) / (nx + 1))) / (dx * dx)
    ey = 2.0 * (1.0 - np.cos(np.pi * np.arange(1, ny + 1, dtype=np.float64) / (ny + 1))) / (dy * dy)
    inv = 1.0 / (ey[:, None] + ex[None, :])

    def apply(flat):
        r = flat.reshape(ny, nx)
        fr = np.where(fluid, r, 0.0)
        t = dstn(fr, type=1, norm="ortho")
        z = np.asarray(idstn(t * inv, type=1, norm="ortho"), dtype=np.float64)
        z[~fluid] = r[~fluid]
        return z.ravel()

    return apply


def solve_masked_poisson(rhs, fluid, dx, dy, periodic_axes=(), tolerance=1.0e-5, max_iterations=120):
    """fb/core/_scipy_adapter.py:16-161. Returns (pressure, info, iterations, relative residual)."""
    ny, nx = rhs.shape
    size = nx * ny
    diag, apply = masked_operator(fluid, dx, dy, periodic_axes)
    op = LinearOperator((size, size), matvec=apply, dtype=rhs.dtype)
    if not periodic_axes:
        pre = dst_preconditioner(fluid, dx, dy)
    else:
        invd = 1.0 / np.maximum(diag, np.finfo(rhs.dtype).eps)

        def pre(flat):
            return invd.ravel() * flat
    M = LinearOperator((size, size), matvec=pre, dtype=rhs.dtype)
    b = rhs.copy()
    singular = "x" in periodic_axes and "y" in periodic_axes
    if singular and np.any(fluid):
        b[fluid] -= np.mean(b[fluid])
    b[~fluid] = 0.0
    if not np.isfinite(b).all():
        raise OracleFailure("nonfinite_state", "pressure RHS contains non-finite values", "projection")
    safe = 0.25 * np.sqrt(np.finfo(rhs.dtype).max / max(size, 1))
    if float(np.max(np.abs(b))) > safe:
        raise OracleFailure("projection_failure", "pressure RHS exceeds the safe CG dot-product range", "projection")
    count = [0]

    def cb(xk):
        count[0] += 1
        if not np.isfinite(xk).all():
            raise OracleFailure("nonfinite_state", "pressure CG produced a non-finite iterate", "projection")

    try:
        with np.errstate(over="raise", invalid="raise", divide="raise"):
            x, info = cg(op, b.ravel(), M=M, rtol=0.9 * tolerance, atol=0.0, maxiter=max_iterations, callback=cb)
    except FloatingPointError as e:
        raise OracleFailure("projection_failure", f"pressure CG failed: {e}", "projection") from e
    p = x.reshape(ny, nx)
    if singular and np.any(fluid):
        p[fluid] -= np.mean(p[fluid])
    res = b.ravel() - apply(x)
    bn = float(np.linalg.norm(b.ravel()))
    rel = float(np.linalg.norm(res)) / max(bn, float(np.finfo(rhs.dtype).eps))
    return p, int(info), count[0], rel


def project_faces(u, v, dom, solid, wall, freestream, dt, channel=False, tol=1.0e-5, maxit=640):
    """fb/core/grid.py:134-180. Returns (u, v, (iterations, residual, converged), pressure)."""
    apply_domain_boundaries(u, v, dom, freestream, channel)
    enforce_solid_faces(u, v, solid, wall)
    div = (u[:, 1:] - u[:, :-1]) / dom.dx + (v[1:, :] - v[:-1, :]) / dom.dy
    fluid = ~solid
    rhs = np.where(fluid, -div / max(dt, 1.0e-12), 0.0)
    p, info, its, rel = solve_masked_poisson(np.asarray(rhs, dtype=np.float64), fluid, dom.dx, dom.dy,
                                            dom.periodic_axes, tol, maxit)
    u[:, 1:-1] -= dt * (p[:, 1:] - p[:, :-1]) / dom.dx
    v[1:-1, :] -= dt * (p[1:, :] - p[:-1, :]) / dom.dy
    if "x" in dom.periodic_axes:
        u[:, 0] -= dt * (p[:, 0] - p[:, -1]) / dom.dx
        u[:, -1] = u[:, 0]
    if "y" in dom.periodic_axes:
        v[0, :] -= dt * (p[0, :] - p[-1, :]) / dom.dy
        v[-1, :] = v[0, :]
    apply_domain_boundaries(u, v, dom, freestream, channel)
    enforce_solid_faces(u, v, solid, wall)
    converged = info == 0 and rel <= tol * (1.0 + 1.0e-6)
    return u, v, (its, rel, converged, info), p


# ----------------------------------------------------------------------------- diffusion

def implicit_diffuse_scalar(field, nu, dt, dom, tol=1.0e-5, maxit=640):
    """fb/core/grid.py:219-275 (Jacobi with edge replication or periodic wrap)."""
    if nu <= 0.0:
        return field.copy(), 0, 0.0, True
    px = "x" in dom.periodic_axes
    py = "y" in dom.periodic_axes
    dupx = px and field.shape[1] == dom.nx + 1
    dupy = py and field.shape[0] == dom.ny + 1
    logical = field[: field.shape[0] - int(dupy), : field.shape[1] - int(dupx)]
    ax = nu * dt / (dom.dx * dom.dx)
    ay = nu * dt / (dom.dy * dom.dy)
    den = 1.0 + 2.0 * ax + 2.0 * ay
    orig = logical.copy()
    res = logical.copy()
    scale = max(float(np.max(np.abs(orig))), 1.0)
    final = float("inf")
    done = 0
    for it in range(1, maxit + 1):
        done = it
        left = np.roll(res, 1, axis=1) if px else np.concatenate((res[:, :1], res[:, :-1]), axis=1)
        right = np.roll(res, -1, axis=1) if px else np.concatenate((res[:, 1:], res[:, -1:]), axis=1)
        low = np.roll(res, 1, axis=0) if py else np.concatenate((res[:1, :], res[:-1, :]), axis=0)
        up = np.roll(res, -1, axis=0) if py else np.concatenate((res[1:, :], res[-1:, :]), axis=0)
        upd = (orig + ax * (left + right) + ay * (low + up)) / den
        final = float(np.max(np.abs(upd - res))) / scale
        res = upd
        if final <= tol:
            break
    conv = final <= tol
    out = field.copy()
    out[: res.shape[0], : res.shape[1]] = res
    if dupx:
        out[:, -1] = out[:, 0]
    if dupy:
        out[-1, :] = out[0, :]
    return np.asarray(out, dtype=field.dtype), done, final, conv


def implicit_diffuse(vel, nu, dt, dom, tol=1.0e-5, maxit=640):
    """fb/core/grid.py:183-216."""
    comps = []
    its = 0
    resid = 0.0
    conv = True
    for k in range(vel.shape[2]):
        c, i, r, ok = implicit_diffuse_scalar(vel[:, :, k], nu, dt, dom, tol, maxit)
        comps.append(c)
        its = max(its, i)
        resid = max(resid, r)
        conv = conv and ok
    return np.stack(comps, axis=2), (its, resid, conv)


# ----------------------------------------------------------------------------- advection

def sample_staggered_scalar(field, points, dom, offset):
    """fb/core/interpolation.py:9-52."""
    x0 = dom.bounds[0][0]
    y0 = dom.bounds[1][0]
    gx = (points[:, 0] - x0) / dom.dx - offset[0]
    gy = (points[:, 1] - y0) / dom.dy - offset[1]
    px = "x" in dom.periodic_axes
    py = "y" in dom.periodic_axes
    gx = np.mod(gx, dom.nx) if px else np.clip(gx, 0.0, field.shape[1] - 1.0)
    gy = np.mod(gy, dom.ny) if py else np.clip(gy, 0.0, field.shape[0] - 1.0)
    ix0 = np.floor(gx).astype(np.int64)
    iy0 = np.floor(gy).astype(np.int64)
    tx = gx - ix0
    ty = gy - iy0
    ix1 = np.mod(ix0 + 1, dom.nx) if px else np.minimum(ix0 + 1, field.shape[1] - 1)
    iy1 = np.mod(iy0 + 1, dom.ny) if py else np.minimum(iy0 + 1, field.shape[0] - 1)
    s = ((1.0 - tx) * (1.0 - ty) * field[iy0, ix0] + tx * (1.0 - ty) * field[iy0, ix1]
         + (1.0 - tx) * ty * field[iy1, ix0] + tx * ty * field[iy1, ix1])
    return np.asarray(s, dtype=field.dtype)


def sample_vector(field, points, dom):
    """fb/core/interpolation.py:55-79."""
    return np.stack([sample_staggered_scalar(field[:, :, k], points, dom, (0.5, 0.5))
                     for k in range(field.shape[2])], axis=1)


def rk2_backtrace(vel, points, dt, dom):
    """fb/core/grid.py:527-537."""
    v0 = sample_vector(vel, points, dom)
    mid = points - 0.5 * dt * v0
    vm = sample_vector(vel, mid, dom)
    return points - dt * vm


def local_velocity_bounds(vel, dom):
    """fb/core/grid.py:540-569."""
    ny, nx, _ = vel.shape
    rows = np.arange(ny)
    cols = np.arange(nx)
    lo = vel.copy()
    hi = vel.copy()
    for oy in (-1, 0, 1):
        r = np.mod(rows + oy, ny) if "y" in dom.periodic_axes else np.clip(rows + oy, 0, ny - 1)
        for ox in (-1, 0, 1):
            c = np.mod(cols + ox, nx) if "x" in dom.periodic_axes else np.clip(cols + ox, 0, nx - 1)
            nb = vel[r[:, None], c[None, :], :]
            lo = np.minimum(lo, nb)
            hi = np.maximum(hi, nb)
    return lo, hi


def advect_velocity(vel, dt, dom, maccormack=True):
    """fb/core/grid.py:506-524."""
    pos = cell_centers(dom.bounds, dom.nx, dom.ny).reshape(-1, 2)
    dep = rk2_backtrace(vel, pos, dt, dom)
    first = sample_vector(vel, dep, dom).reshape(vel.shape)
    if not maccormack:
        return first
    fp = rk2_backtrace(first, pos, -dt, dom)
    fwd = sample_vector(first, fp, dom).reshape(vel.shape)
    corr = first + 0.5 * (vel - fwd)
    lo, hi = local_velocity_bounds(vel, dom)
    lo = np.minimum(lo, first)
    hi = np.maximum(hi, first)
    return np.asarray(np.clip(corr, lo, hi), dtype=vel.dtype)


# ----------------------------------------------------------------------------- face-native transports

def implicit_diffuse_faces(u, v, nu, dt, dom, tol=1.0e-5, maxit=640):
    """fb/core/grid.py:278-303: the scalar Jacobi solve on each face component."""
    ou, iu, ru, cu = implicit_diffuse_scalar(u, nu, dt, dom, tol, maxit)
    ov, iv, rv, cv = implicit_diffuse_scalar(v, nu, dt, dom, tol, maxit)
    return ou, ov, (max(iu, iv), max(ru, rv), cu and cv)


def face_positions(dom):
    """fb/core/grid.py:306-320: u-face points (ny*(nx+1), 2) and v-face points ((ny+1)*nx, 2)."""
    (x0, x1), (y0, y1) = dom.bounds
    ux = np.linspace(x0, x1, dom.nx + 1)
    uy = np.linspace(y0 + 0.5 * dom.dy, y1 - 0.5 * dom.dy, dom.ny)
    vx = np.linspace(x0 + 0.5 * dom.dx, x1 - 0.5 * dom.dx, dom.nx)
    vy = np.linspace(y0, y1, dom.ny + 1)
    pu = np.stack(np.meshgrid(ux, uy), axis=2).reshape(-1, 2)
    pv = np.stack(np.meshgrid(vx, vy), axis=2).reshape(-1, 2)
    return pu, pv


def face_local_bounds(f):
    """fb/core/grid.py:335-347: min and max over each point's 3x3 neighbourhood, edges replicated."""
    h, w = f.shape
    padded = np.pad(f, 1, mode="edge")
    around = [padded[1 + oy: 1 + oy + h, 1 + ox: 1 + ox + w] for oy in (-1, 0, 1) for ox in (-1, 0, 1)]
    return np.minimum.reduce(around), np.maximum.reduce(around)


def _face_departure_sample(f, vel, points, offset, dt, dom):
    """fb/core/grid.py:323-332: the component sampled at its faces' RK2 departure points."""
    return sample_staggered_scalar(f, rk2_backtrace(vel, points, dt, dom), dom, offset).reshape(f.shape)


def advect_faces(u, v, dt, dom, maccormack=True):
    """fb/core/grid.py:350-376: semi-Lagrangian advection of each face component through the
    cell velocity; with MacCormack, the backward-forward correction clipped to the component's
    own 3x3 bounds."""
    vel = faces_to_cell(u, v)
    pu, pv = face_positions(dom)
    first_u = _face_departure_sample(u, vel, pu, (0.0, 0.5), dt, dom)
    first_v = _face_departure_sample(v, vel, pv, (0.5, 0.0), dt, dom)
    if not maccormack:
        return first_u, first_v
    fvel = faces_to_cell(first_u, first_v)
    back_u = _face_departure_sample(first_u, fvel, pu, (0.0, 0.5), -dt, dom)
    back_v = _face_departure_sample(first_v, fvel, pv, (0.5, 0.0), -dt, dom)
    corr_u = first_u + 0.5 * (u - back_u)
    corr_v = first_v + 0.5 * (v - back_v)
    lo, hi = face_local_bounds(u)
    corr_u = np.clip(corr_u, lo, hi)
    lo, hi = face_local_bounds(v)
    corr_v = np.clip(corr_v, lo, hi)
    return np.asarray(corr_u, dtype=u.dtype), np.asarray(corr_v, dtype=v.dtype)


def centered_derivative(f, h, axis, periodic, duplicate_endpoint=False):
    """fb/core/grid.py:379-405: central differences with wrap (over the logical points when the
    last one duplicates the first), else numpy.gradient with second-order one-sided edges."""
    if periodic:
        if duplicate_endpoint:
            n = f.shape[axis] - 1
            core = np.take(f, np.arange(n), axis=axis)
            d = (np.roll(core, -1, axis=axis) - np.roll(core, 1, axis=axis)) / (2.0 * h)
            return np.asarray(np.concatenate((d, np.take(d, [0], axis=axis)), axis=axis), dtype=f.dtype)
        return np.asarray((np.roll(f, -1, axis=axis) - np.roll(f, 1, axis=axis)) / (2.0 * h), dtype=f.dtype)
    return np.asarray(np.gradient(f, h, axis=axis, edge_order=2), dtype=f.dtype)


def cross_velocity_on_faces(u, v, dom):
    """fb/core/grid.py:408-431: v averaged onto u faces and u onto v faces through the cells."""
    cu = 0.5 * (u[:, :-1] + u[:, 1:])
    cv = 0.5 * (v[:-1, :] + v[1:, :])
    v_on_u = np.empty_like(u)
    v_on_u[:, 1:-1] = 0.5 * (cv[:, :-1] + cv[:, 1:])
    if "x" in dom.periodic_axes:
        edge = 0.5 * (cv[:, -1] + cv[:, 0])
        v_on_u[:, 0] = edge
        v_on_u[:, -1] = edge
    else:
        v_on_u[:, 0] = cv[:, 0]
        v_on_u[:, -1] = cv[:, -1]
    u_on_v = np.empty_like(v)
    u_on_v[1:-1, :] = 0.5 * (cu[:-1, :] + cu[1:, :])
    if "y" in dom.periodic_axes:
        edge = 0.5 * (cu[-1, :] + cu[0, :])
        u_on_v[0, :] = edge
        u_on_v[-1, :] = edge
    else:
        u_on_v[0, :] = cu[0, :]
        u_on_v[-1, :] = cu[-1, :]
    return v_on_u, u_on_v


def skew_face_advection_rate(u, v, dom):
    """fb/core/grid.py:434-443: max over faces of |u_n|/dx + |u_t|/dy (u faces) and the v-face
    counterpart."""
    v_on_u, u_on_v = cross_velocity_on_faces(u, v, dom)
    ru = np.abs(u) / dom.dx + np.abs(v_on_u) / dom.dy
    rv = np.abs(u_on_v) / dom.dx + np.abs(v) / dom.dy
    return max(float(np.max(ru)), float(np.max(rv)))


def skew_symmetric_convection(u, v, dom):
    """fb/core/grid.py:446-478: half the advective plus half the conservative convection."""
    v_on_u, u_on_v = cross_velocity_on_faces(u, v, dom)
    px = "x" in dom.periodic_axes
    py = "y" in dom.periodic_axes
    d = centered_derivative
    adv_u = u * d(u, dom.dx, 1, px, px) + v_on_u * d(u, dom.dy, 0, py)
    adv_v = u_on_v * d(v, dom.dx, 1, px) + v * d(v, dom.dy, 0, py, py)
    con_u = d(u * u, dom.dx, 1, px, px) + d(v_on_u * u, dom.dy, 0, py)
    con_v = d(u_on_v * v, dom.dx, 1, px) + d(v * v, dom.dy, 0, py, py)
    return np.asarray(0.5 * (adv_u + con_u), dtype=u.dtype), np.asarray(0.5 * (adv_v + con_v), dtype=v.dtype)


def advect_faces_skew_rk2(u, v, dt, dom, solid, wall, freestream):
    """fb/core/grid.py:481-503: explicit midpoint step; boundaries (without channel walls) and
    solid faces are imposed on the midpoint and on the result."""
    ku, kv = skew_symmetric_convection(u, v, dom)
    mid_u = np.asarray(u - 0.5 * dt * ku, dtype=u.dtype)
    mid_v = np.asarray(v - 0.5 * dt * kv, dtype=v.dtype)
    apply_domain_boundaries(mid_u, mid_v, dom, freestream)
    enforce_solid_faces(mid_u, mid_v, solid, wall)
    ku, kv = skew_symmetric_convection(mid_u, mid_v, dom)
    out_u = np.asarray(u - dt * ku, dtype=u.dtype)
    out_v = np.asarray(v - dt * kv, dtype=v.dtype)
    apply_domain_boundaries(out_u, out_v, dom, freestream)
    enforce_solid_faces(out_u, out_v, solid, wall)
    return out_u, out_v


# ----------------------------------------------------------------------------- solver

def _opt(sc, name, default):
    return sc.solver_options.get(name, default)


def initial_velocity(sc):
    """fb/solvers/stable_fluids.py:102-116."""
    dom = sc.domain
    vel = np.empty((dom.ny, dom.nx, 2), dtype=sc.dtype)
    vel[...] = np.asarray(sc.freestream[:2], dtype=sc.dtype)
    init = str(_opt(sc, "initial_condition", "freestream"))
    pos = cell_centers(dom.bounds, dom.nx, dom.ny)
    if init == "taylor-green":
        vel[:, :, 0] = np.sin(pos[:, :, 0]) * np.cos(pos[:, :, 1])
        vel[:, :, 1] = -np.cos(pos[:, :, 0]) * np.sin(pos[:, :, 1])
    elif init == "poiseuille":
        y0, y1 = dom.bounds[1]
        r = 0.5 * (y1 - y0)
        c = 0.5 * (y0 + y1)
        vel[:, :, 0] = 1.5 * (1.0 - ((pos[:, :, 1] - c) / r) ** 2)
        vel[:, :, 1] = 0.0
    return vel


class StableFluidsOracle:
    """Restates ``StableFluidsSolver`` (fb/solvers/stable_fluids.py:119-603): cell-centred
    MacCormack or semi-Lagrangian transport, face-native transport (``stable_face_advection``)
    and skew-symmetric RK2 (``stable_advection="skew-rk2"``)."""

    def initialize(self, sc, seed=0, solid_fn=None):
        # fb/solvers/stable_fluids.py:195-216
        self.sc = sc
        self.dom = Dom.of(sc.domain)
        self.foil = NacaFoil(sc.foil.naca, sc.foil.chord, tuple(sc.foil.pivot[:2]))
        self.solid_fn = solid_fn or (lambda angle: self.foil.mask(self.dom.bounds, self.dom.nx, self.dom.ny, angle))
        self.control = sc.control_at(0.0)
        self.time = 0.0
        self.revision = 0
        self.reynolds = float(sc.reynolds)
        mode = str(_opt(sc, "stable_advection", "maccormack"))
        self.maccormack = mode == "maccormack"
        self.skew = mode == "skew-rk2"
        self.face_advection = bool(_opt(sc, "stable_face_advection", False))
        self.u, self.v = cell_to_faces(initial_velocity(sc))
        self.solid = self.solid_fn(self.control.angle_degrees)
        self.last_projection = (0, 0.0, True, 0)
        self.last_viscosity = (0, 0.0, True)
        self.apply_projection(max(sc.output_dt, 1.0e-4))
        self.revision = 0

    def wall_grid(self, angular_velocity_degrees):
        pts = cell_centers(self.dom.bounds, self.dom.nx, self.dom.ny).reshape(-1, 2)
        return self.foil.wall_velocity(pts, angular_velocity_degrees).reshape(self.dom.ny, self.dom.nx, 2)

    def apply_projection(self, dt):
        # fb/solvers/stable_fluids.py:256-330
        sc = self.sc
        dom = self.dom
        channel = str(_opt(sc, "initial_condition", "")) == "poiseuille"
        tol = float(_opt(sc, "pressure_tolerance", 1.0e-5))
        maxit = int(_opt(sc, "pressure_max_iterations", 640))
        cfl = float(_opt(sc, "stable_cfl", 0.7))
        if self.skew:
            cfl = min(cfl, 0.4)
        limit = max(1.0, 2.0 * cfl)
        wall = self.wall_grid(self.control.angular_velocity_degrees)
        if not (np.isfinite(self.u).all() and np.isfinite(self.v).all()):
            raise OracleFailure("nonfinite_state", "projection received non-finite velocity", "projection")
        face = max(float(np.max(np.abs(self.u))), float(np.max(np.abs(self.v))))
        ws = float(np.max(np.linalg.norm(wall[self.solid], axis=1))) if np.any(self.solid) else 0.0
        pc = max(face, ws) * dt / min(dom.dx, dom.dy)
        if pc > limit:
            raise OracleFailure("excessive_velocity", "projection CFL", "projection",
                                {"maximum_cfl": pc, "maximum_cfl_limit": limit})
        self.u, self.v, rep, self.pressure = project_faces(self.u, self.v, dom, self.solid, wall, sc.freestream, dt,
                                                          channel, tol, maxit)
        self.last_projection = rep
        if not rep[2]:
            raise OracleFailure("projection_failure", "pressure CG did not converge", "projection",
                                {"iterations": rep[0], "relative_residual": rep[1]})

    def cell_velocity(self):
        return faces_to_cell(self.u, self.v)

    def import_state(self, velocity, time, control):
        # fb/solvers/stable_fluids.py:636-706: cell velocity with the solid cells at the wall velocity of the
        # import pose, to faces, projected with a 1e-4 pseudo-step; rolled back on failure
        saved = (self.u.copy(), self.v.copy(), self.solid.copy(), self.control, self.time, self.revision,
                 self.last_projection)
        try:
            vel = np.asarray(velocity, dtype=self.sc.dtype).copy()
            self.time = float(time)
            self.control = control
            self.solid = self.solid_fn(control.angle_degrees)
            wall = self.wall_grid(control.angular_velocity_degrees)
            vel[self.solid] = wall[self.solid]
            self.u, self.v = cell_to_faces(vel)
            self.apply_projection(1.0e-4)
        except Exception:
            (u, v, sol, self.control, self.time, self.revision, self.last_projection) = saved
            self.u, self.v, self.solid = u, v, sol
            raise
        self.revision += 1

    def advance(self, control, target_dt):
        # fb/solvers/stable_fluids.py:332-603
        sc = self.sc
        dom = self.dom
        cv = self.cell_velocity()
        max_speed = max(float(np.max(np.linalg.norm(cv, axis=2))), abs(sc.freestream[0]), 1.0e-6)
        cfl = float(_opt(sc, "stable_cfl", 0.7))
        if self.skew:
            cfl = min(cfl, 0.4)
        spacing = min(dom.dx, dom.dy)
        wall_speed = abs(np.deg2rad(control.angular_velocity_degrees)) * self.foil.maximum_radius
        sweep = abs(np.deg2rad(control.angle_degrees - self.control.angle_degrees)) * self.foil.maximum_radius / spacing
        rate = skew_face_advection_rate(self.u, self.v, dom) if self.skew else max_speed / spacing
        measure = target_dt * rate
        required = max(1.05 * measure / cfl, target_dt * wall_speed / (cfl * spacing), sweep / cfl)
        subs = max(1, int(np.ceil(required)))
        if subs > 512:
            raise OracleFailure("stability_limit", "too many substeps", "boundary", {"required_substeps": subs})
        dt = target_dt / subs
        nu = sc.reference_speed * sc.foil.chord / self.reynolds
        tol = float(_opt(sc, "pressure_tolerance", 1.0e-5))
        maxit = int(_opt(sc, "pressure_max_iterations", 640))
        saved = (self.u.copy(), self.v.copy(), self.solid.copy(), self.control, self.time, self.revision,
                 self.last_projection, self.last_viscosity)
        start_time = self.time
        start_angle = self.control.angle_degrees
        retries = 0

        def restore():
            (u, v, s, self.control, self.time, self.revision, self.last_projection, self.last_viscosity) = saved
            self.u, self.v, self.solid = u.copy(), v.copy(), s.copy()

        while True:
            try:
                for k in range(subs):
                    frac = (k + 1) / subs
                    sub = type(control)(start_time + frac * target_dt,
                                        start_angle + frac * (control.angle_degrees - start_angle),
                                        control.angular_velocity_degrees)
                    if self.skew:
                        wall = self.wall_grid(sub.angular_velocity_degrees)
                        au, av = advect_faces_skew_rk2(self.u, self.v, dt, dom, self.solid, wall, sc.freestream)
                        self.u, self.v, self.last_viscosity = implicit_diffuse_faces(au, av, nu, dt, dom, tol, maxit)
                    elif self.face_advection:
                        au, av = advect_faces(self.u, self.v, dt, dom, self.maccormack)
                        self.u, self.v, self.last_viscosity = implicit_diffuse_faces(au, av, nu, dt, dom, tol, maxit)
                    else:
                        vel = faces_to_cell(self.u, self.v)
                        vel = advect_velocity(vel, dt, dom, self.maccormack)
                        vel, self.last_viscosity = implicit_diffuse(vel, nu, dt, dom, tol, maxit)
                        self.u, self.v = cell_to_faces(vel)
                    if not self.last_viscosity[2]:
                        raise OracleFailure("convergence_failure", "implicit viscosity did not converge", "viscosity")
                    self.control = sub
                    self.solid = self.solid_fn(sub.angle_degrees)
                    self.apply_projection(dt)
                fv = self.cell_velocity()
                final_speed = float(np.max(np.linalg.norm(fv, axis=2)))
                accepted = dt * (skew_face_advection_rate(self.u, self.v, dom) if self.skew else final_speed / spacing)
            except Exception:
                restore()
                raise
            if accepted <= cfl * (1.0 + 1.0e-6):
                break
            nxt = max(subs + 1, int(np.ceil(1.05 * subs * accepted / cfl)))
            restore()
            if nxt > 512:
                raise OracleFailure("stability_limit", "retry requires too many substeps", "advection")
            retries += 1
            subs = nxt
            dt = target_dt / subs
        wall = self.wall_grid(control.angular_velocity_degrees)
        div = native_divergence_linf(self.u, self.v, dom, self.solid)
        leak = solid_face_leakage(self.u, self.v, self.solid, wall)
        self.time = start_time + target_dt
        self.control = type(control)(self.time, control.angle_degrees, control.angular_velocity_degrees)
        self.revision += 1
        return {"substeps": subs, "max_speed": final_speed, "stability_retries": retries,
                "pressure_iterations": self.last_projection[0],
                "pressure_relative_residual": self.last_projection[1],
                "viscosity_iterations": self.last_viscosity[0],
                "viscosity_final_residual": self.last_viscosity[1],
                "divergence_linf": div, "solid_leakage": leak}
