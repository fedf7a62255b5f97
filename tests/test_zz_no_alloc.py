"""fb_pic_particle_to_faces_noalloc (P2G with its sort buffers in the library's fixed device array) against
fb_pic_particle_to_faces (the same transfer on host arrays, with its own allocations): identical faces and
unsupported-face counts, bit for bit, on two calls in a row of which the second is smaller, for walled and
periodic boxes in FP64 and FP32; inputs beyond the fixed maxima are rejected before any work is queued.

Device buffers for the inputs and outputs come from hipMalloc of the HIP runtime the library itself links (the
test's plumbing: the entry point under test allocates nothing)."""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FB_F64, FB_F32 = 0, 1
H2D, D2H = 1, 2


class Hip:
    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(int(nbytes), 16)) == 0
        self.bufs.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
        return p

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _case(rng, nx, ny, n, dtype, periodic):
    x0, y0, dx, dy = -1.0, -0.5, 2.0 / nx, 1.0 / ny
    pos = np.empty((n, 2), dtype)
    pos[:, 0] = x0 + nx * dx * rng.random(n)
    pos[:, 1] = y0 + ny * dy * rng.random(n)
    # sorted by cell in part, as seeded particles are, and scattered in part
    k = n // 2
    order = np.lexsort((pos[:k, 0], np.floor((pos[:k, 1] - y0) / dy)))
    pos[:k] = pos[:k][order]
    vel = rng.standard_normal((n, 2)).astype(dtype)
    fu = rng.standard_normal((ny, nx + 1)).astype(dtype)
    fv = rng.standard_normal((ny + 1, nx)).astype(dtype)
    return dict(nx=nx, ny=ny, x0=x0, y0=y0, dx=dx, dy=dy, pos=pos, vel=vel, fu=fu, fv=fv, px=int(periodic),
                py=int(periodic))


def _reference(ctx, prec, c):
    from foilbench_amd import _native as N

    dt = c["pos"].dtype
    u = np.empty((c["ny"], c["nx"] + 1), dt)
    v = np.empty((c["ny"] + 1, c["nx"]), dt)
    frac = C.c_double()
    N.check(ctx.lib.fb_pic_particle_to_faces(ctx.handle, prec, c["pos"].shape[0], N.ptr(c["pos"]), N.ptr(c["vel"]),
                                             N.ptr(c["fu"]), N.ptr(c["fv"]), c["nx"], c["ny"], c["x0"], c["y0"],
                                             c["dx"], c["dy"], c["px"], c["py"], N.ptr(u), N.ptr(v), C.byref(frac)))
    return u, v, frac.value


def _noalloc(ctx, hip, prec, c):
    from foilbench_amd import _native as N

    dt = c["pos"].dtype
    dp, dv, dfu, dfv = hip.put(c["pos"]), hip.put(c["vel"]), hip.put(c["fu"]), hip.put(c["fv"])
    du = hip.alloc(c["ny"] * (c["nx"] + 1) * dt.itemsize)
    dvv = hip.alloc((c["ny"] + 1) * c["nx"] * dt.itemsize)
    dk = hip.put(np.full(1, 12345, np.uint64))  # stale value: the call must clear it
    N.check(ctx.lib.fb_pic_particle_to_faces_noalloc(ctx.handle, prec, c["pos"].shape[0], dp, dv, dfu, dfv, c["nx"],
                                                     c["ny"], c["x0"], c["y0"], c["dx"], c["dy"], c["px"], c["py"],
                                                     du, dvv, dk))
    ctx.synchronize()
    u = hip.get(du, (c["ny"], c["nx"] + 1), dt)
    v = hip.get(dvv, (c["ny"] + 1, c["nx"]), dt)
    k = int(hip.get(dk, (1,), np.uint64)[0])
    return u, v, k / (u.size + v.size)


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("dtype,prec", [(np.float64, FB_F64), (np.float32, FB_F32)])
def test_noalloc_p2g_equals_allocating_entry_point(dtype, prec, periodic):
    from foilbench_amd import _native as N

    ctx = N.Context.default()
    hip = Hip()
    rng = np.random.default_rng(7)
    try:
        # two calls in a row, the second smaller (fewer particles, fewer cells): the fixed buffers keep the
        # first call's contents
        for nx, ny, n in ((96, 48, 30000), (40, 20, 2500)):
            c = _case(rng, nx, ny, n, dtype, periodic)
            print(f"noalloc p2g {np.dtype(dtype).name} periodic={periodic} {nx}x{ny} n={n}", flush=True)
            ru, rv, rf = _reference(ctx, prec, c)
            gu, gv, gf = _noalloc(ctx, hip, prec, c)
            np.testing.assert_array_equal(gu, ru)
            np.testing.assert_array_equal(gv, rv)
            assert gf == rf
            assert np.isfinite(gu).all() and np.isfinite(gv).all()
    finally:
        hip.free()


def test_noalloc_p2g_rejects_inputs_beyond_its_maxima():
    from foilbench_amd import _native as N

    ctx = N.Context.default()
    hip = Hip()
    try:
        d = hip.alloc(64)
        lib = ctx.lib
        # too many particles (the maximum is 2^19)
        rc = lib.fb_pic_particle_to_faces_noalloc(ctx.handle, FB_F32, (1 << 19) + 1, d, d, d, d, 64, 32, 0.0, 0.0,
                                                  0.1, 0.1, 0, 0, d, d, d)
        assert rc == 1 and b"allocation-free maximum" in lib.fb_last_error()
        # too many cells (the maximum is 2^18)
        rc = lib.fb_pic_particle_to_faces_noalloc(ctx.handle, FB_F32, 10, d, d, d, d, 1024, 512, 0.0, 0.0, 0.1, 0.1,
                                                  0, 0, d, d, d)
        assert rc == 1 and b"allocation-free maximum" in lib.fb_last_error()
        rc = lib.fb_pic_particle_to_faces_noalloc(ctx.handle, FB_F32, 10, d, d, d, d, 64, 32, 0.0, 0.0, 0.1, 0.1, 0,
                                                  0, d, d, None)
        assert rc == 1
    finally:
        hip.free()
