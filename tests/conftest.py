"""Shared fixtures: the oracle (CPU checker) and the C-ABI binding of the HIP library."""
import importlib.util
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


oracle = _load("shrt_oracle_py", os.path.join(REPO, "oracle", "oracle.py"))
shrt = _load("shrt_binding", os.path.join(REPO, "shearing-rt_amd", "shrt.py"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")


@pytest.fixture(scope="session")
def gpu():
    """Skip when no HIP device is visible; fail if the library is missing."""
    L = shrt.lib()   # raises if libshrt_hip.so was not built
    if L.shrt_device_count() < 1:
        pytest.skip("no HIP device visible")
    return shrt


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def random_spectrum(rng, nkx, nky, decay=2.0, amp=1e-2):
    """Random spectral field with |f_hat| ~ amp / (1 + |index|)^decay."""
    ikx = np.arange(nkx)
    kxi = np.where(ikx < (nkx + 1) // 2, ikx, ikx - nkx)
    kk = np.sqrt(kxi[:, None] ** 2 + np.arange(nky)[None, :] ** 2)
    z = rng.standard_normal((nkx, nky)) + 1j * rng.standard_normal((nkx, nky))
    return amp * z / (1.0 + kk) ** decay
