"""The C-ABI library loads and exports every entry point include/shrt.h declares;
argument validation and size bookkeeping work without touching a GPU."""
import ctypes

import numpy as np
import pytest

from conftest import oracle, shrt


def test_library_exports_every_declared_symbol():
    L = shrt.lib()
    declared = shrt.header_symbols()
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(L, name), name
    # and the binding covers all of them
    assert set(declared) <= set(shrt._SIGS)


def test_version_and_no_error():
    L = shrt.lib()
    assert L.shrt_version().decode().startswith("shrt")
    assert isinstance(L.shrt_last_error(), bytes)


@pytest.mark.parametrize("nx,ny", [(128, 128), (2048, 2048), (64, 256), (4096, 4096), (100, 100), (6, 9)])
def test_sizes_match_oracle(nx, ny):
    a, b = shrt.sizes(nx, ny), oracle.sizes(nx, ny)
    for k in ("nkx", "nky", "nkxh", "nkxh2", "nkxpad", "ncy", "M"):
        assert a[k] == b[k], k
    # the library's intermediate arrays hold the kx >= 0 half: nkxh * ny
    assert a["I"] == a["nkxh"] * ny


def test_supported_sizes():
    L = shrt.lib()
    for n in (32, 64, 128, 256, 512, 1024, 2048):
        assert L.shrt_supported(n, n, shrt.FP32) == 1
        assert L.shrt_supported(n, n, shrt.FP64) == 1
    assert L.shrt_supported(4096, 4096, shrt.FP32) == 1
    assert L.shrt_supported(4096, 4096, shrt.FP64) == 1     # the one-team step kernels (67.6 KB of LDS)
    assert L.shrt_supported(100, 100, shrt.FP32) == 0
    assert L.shrt_supported(128, 128, 3) == 0


def _create_status(**kw):
    p = dict(shrt.Shrt.DEFAULTS)
    p.update(kw)
    ints = ("nx", "ny", "noise", "linear", "precision", "graph_steps")
    prm = shrt.Params(**{k: (int(v) if k in ints else float(v)) for k, v in p.items()})
    h = ctypes.c_void_p()
    st = shrt.lib().shrt_create(ctypes.byref(prm), ctypes.byref(h))
    if st == 0:
        shrt.lib().shrt_destroy(h)
    return st, shrt.lib().shrt_last_error().decode()


@pytest.mark.parametrize("kw,code", [
    (dict(nx=100, ny=100), 2),                    # not a power of two
    (dict(precision=3), 1),
    (dict(graph_steps=3), 1),
    (dict(delt=0.0), 1),
    (dict(nx=256, ny=128, sigma=1.0), 1),         # remap fold would leave the array
    (dict(sigma=500.0, delt=1e-3), 1),            # remaps closer than 3 steps
    (dict(nx=8192, ny=8192, precision=4), 2),     # no 8192 transform compiled in
])
def test_create_rejects_bad_parameters(kw, code):
    st, msg = _create_status(**kw)
    assert st == code
    assert msg


def test_null_arguments():
    L = shrt.lib()
    assert L.shrt_create(None, None) == 1
    assert L.shrt_step(None, 1) == 1
    assert L.shrt_get_spectral(None, 0, None) == 1
    assert L.shrt_destroy(None) == 0
    # entry points that check their arguments before touching the device
    assert L.shrt_get_physical(None, 5, 0, None) == 1
    assert L.shrt_get_variants(None, None, None, None, None) == 1
    assert L.shrt_set_variants(None, 0, 0, 0) == 1
    assert L.shrt_closure_correct(None, None, None, 1.0, None, 0.1, 1.0, 0, None) == 1
    assert L.shrt_copy_bench(0, 1, None) == 1
    assert L.shrt_stream_copy_bench(24, 1, None) == 1
