"""shrt_xtok_noalloc: the allocation-free xtok on device buffers, through the C
ABI, against shrt_xtok on the same inputs (the same kernels, so bit for bit),
two calls in a row with the second on a smaller grid (both use the library's
one scratch array), FP64 as well, and the refusal of a grid whose intermediate
exceeds the scratch."""
import ctypes

import numpy as np
import pytest

from conftest import shrt

pytestmark = pytest.mark.gpu

BASE = dict(Lx=2 * np.pi, Ly=2 * np.pi, nu=1e-3, kappa=1e-3, sigma=0.5, g=1.0, rho0=1.0,
            rho0_prime=1.0, delt=1e-3, cfl_num=0.35, noise=0, rho_eps1=1e-2)


class _Hip:
    """The few HIP runtime calls the test needs for its own device buffers."""

    def __init__(self):
        shrt.lib()   # the library's own HIP runtime, by its soname
        try:
            self.L = ctypes.CDLL("libamdhip64.so.7")
        except OSError:
            self.L = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def h2d(self, dst, arr):
        assert self.L.hipMemcpy(dst, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, 1) == 0

    def d2h(self, arr, src):
        assert self.L.hipMemcpy(arr.ctypes.data_as(ctypes.c_void_p), src, arr.nbytes, 2) == 0

    def free(self, p):
        assert self.L.hipFree(p) == 0


def _noalloc(hip, g, f):
    """phys (host) -> device buffer -> shrt_xtok_noalloc -> host spectrum."""
    x = np.ascontiguousarray(f, dtype=g.real)
    y = np.zeros(g.shape_k, dtype=g.cplx)
    dx, dy = hip.alloc(x.nbytes), hip.alloc(y.nbytes)
    try:
        hip.h2d(dx, x)
        st = shrt.lib().shrt_xtok_noalloc(g.h, dx, dy)
        assert st == 0, shrt.lib().shrt_last_error().decode()
        assert shrt.lib().shrt_synchronize(g.h) == 0
        hip.d2h(y, dy)
    finally:
        hip.free(dx)
        hip.free(dy)
    return y


@pytest.mark.parametrize("prec,sizes", [(4, [(256, 256), (64, 64)]), (8, [(128, 256), (32, 32)])])
def test_xtok_noalloc_matches_xtok(gpu, prec, sizes):
    hip = _Hip()
    rng = np.random.default_rng(11)
    ctxs = []
    for nx, ny in sizes:
        print(f"xtok_noalloc prec={prec} nx={nx} ny={ny}", flush=True)
        g = shrt.Shrt(nx=nx, ny=ny, precision=prec, **BASE)
        ctxs.append(g)
        f = rng.standard_normal((nx, ny))
        want = g.xtok(f)
        got = _noalloc(hip, g, f)
        assert got.shape == want.shape
        assert np.array_equal(got, want)
        # a second call on the same context after another one: same result
        assert np.array_equal(_noalloc(hip, g, f), want)
    # the larger context again after the smaller one wrote the scratch
    f = rng.standard_normal(sizes[0])
    assert np.array_equal(_noalloc(hip, ctxs[0], f), ctxs[0].xtok(f))
    for g in ctxs:
        g.close()


def test_xtok_noalloc_refuses_oversize(gpu):
    g = shrt.Shrt(nx=2048, ny=2048, precision=8, **BASE)
    st = shrt.lib().shrt_xtok_noalloc(g.h, ctypes.c_void_p(1), ctypes.c_void_p(1))
    assert st == 1
    assert b"scratch" in shrt.lib().shrt_last_error()
    g.close()
