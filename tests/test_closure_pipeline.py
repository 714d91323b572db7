"""Host logic of the closure pipeline (shearing-rt_amd/closure.py): the
fine-to-coarse mode filter keeps exactly the coarse grid's (kx, ky) and the
enstrophy spectrum sums the right shells with the Hermitian weights."""
import importlib.util
import os

import numpy as np

from conftest import shrt

HERE = os.path.dirname(os.path.abspath(__file__))
spec = importlib.util.spec_from_file_location("closure_pipe", os.path.join(HERE, "..", "shearing-rt_amd", "closure.py"))
cl = importlib.util.module_from_spec(spec)
spec.loader.exec_module(cl)


def test_filter_keeps_the_coarse_modes():
    for nf, nc in [(128, 32), (64, 32), (256, 64)]:
        sf, sc = shrt.sizes(nf, nf), shrt.sizes(nc, nc)
        kxf, _ = cl.k_ints(sf)
        kxc, kyc = cl.k_ints(sc)
        rows = cl.coarse_rows(sf, sc)
        assert np.array_equal(kxf[rows], kxc)
        F = (kxf[:, None] * 1000 + np.arange(sf["nky"])[None, :]).astype(np.complex128)
        G = cl.Filter(sf, sc)(F)
        assert G.shape == (sc["nkx"], sc["nky"])
        assert np.array_equal(G.real, kxc[:, None] * 1000 + kyc[None, :])


def test_enstrophy_spectrum_shells():
    s = shrt.sizes(32, 32)
    kx, ky = cl.k_ints(s)
    om = np.zeros((s["nkx"], s["nky"]), dtype=np.complex128)
    om[3, 4] = 2.0          # |k| = 5, ky > 0: counted twice
    om[0, 0] = 1.0          # mean: shell 0, once
    row = int(np.where(kx == -6)[0][0])
    om[row, 0] = 1j         # |k| = 6, ky = 0: once
    Z = cl.enstrophy_spectrum(om, kx, ky)
    assert Z[5] == 0.5 * 2 * 4.0 and Z[0] == 0.5 and Z[6] == 0.5
    assert Z.sum() == 4.0 + 0.5 + 0.5
