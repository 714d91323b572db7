"""Pins the oracle's FFT, xtok and ktox against independent references:
brute-force DFT sums, the committed numpy.fft (pocketfft) golden vectors, Parseval,
round trips and a single-mode known answer (cos(2x + 10y), the reference IC shape)."""
import os

import numpy as np
import pytest

from conftest import oracle

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz"))


def brute_dft(x, sign):
    n = len(x)
    k = np.arange(n)
    return np.array([np.sum(x * np.exp(sign * 2j * np.pi * k * j / n)) for j in range(n)])


@pytest.mark.parametrize("n", [2, 4, 8, 16, 32, 6, 12])
def test_fft_brute_force(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(oracle.fft(x, -1), brute_dft(x, -1), atol=1e-12)
    assert np.allclose(oracle.fft(x, +1), brute_dft(x, +1), atol=1e-12)


@pytest.mark.parametrize("n", [8, 16, 32, 64, 128, 256, 12])
def test_fft_golden(n):
    x = GOLD[f"c{n}_in"]
    np.testing.assert_allclose(oracle.fft(x, -1), GOLD[f"c{n}_fwd"], rtol=0, atol=1e-12 * n)
    np.testing.assert_allclose(oracle.fft(x, +1), GOLD[f"c{n}_inv"], rtol=0, atol=1e-12 * n)


def _rows(s):
    return list(range(s["nkxh"])) + [i + s["nkxpad"] for i in range(s["nkxh"], s["nkx"])]


@pytest.mark.parametrize("nx,ny", [(32, 48), (64, 64)])
def test_xtok_golden(nx, ny):
    o = oracle.Oracle(nx=nx, ny=ny, Lx=2 * np.pi, Ly=2 * np.pi)
    s = o.s
    F = o.xtok(GOLD[f"r{nx}x{ny}_in"])
    ref = GOLD[f"r{nx}x{ny}_rfft2"][_rows(s), : s["nky"]] / (nx * ny)
    np.testing.assert_allclose(F, ref, rtol=0, atol=1e-15)


@pytest.mark.parametrize("nx,ny", [(32, 48), (64, 64)])
def test_ktox_golden(nx, ny):
    # ktox of the retained part equals irfft2 of the zero-padded half spectrum
    o = oracle.Oracle(nx=nx, ny=ny, Lx=2 * np.pi, Ly=2 * np.pi)
    s = o.s
    P = GOLD[f"k{nx}x{ny}_in"]
    F = P[_rows(s), : s["nky"]]
    Pz = np.zeros_like(P)
    Pz[_rows(s), : s["nky"]] = F
    ref = np.fft.irfft2(Pz, s=(nx, ny)) * nx * ny
    np.testing.assert_allclose(o.ktox(F), ref, rtol=0, atol=1e-10)


def test_dealias_boundaries():
    """Row nkxh-1 and column nky-1 are kept, row nkxh (|kx| = nkxh) is dropped."""
    n = 48
    o = oracle.Oracle(nx=n, ny=n, Lx=2 * np.pi, Ly=2 * np.pi)
    s = o.s
    x = np.arange(n) * 2 * np.pi / n
    for kx, ky, kept in [(s["nkxh"] - 1, 0, True), (s["nkxh"], 0, False), (0, s["nky"] - 1, True),
                         (0, s["nky"], False)]:
        f = np.cos(kx * x[:, None] + ky * x[None, :])
        F = o.xtok(f)
        assert (np.abs(F).max() > 0.4) == kept, (kx, ky)


def test_single_mode_known_answer():
    """The reference IC rho = eps*cos(2x + 10y) has two stored modes of eps/2 (one stored)."""
    n = 64
    o = oracle.Oracle(nx=n, ny=n, Lx=2 * np.pi, Ly=2 * np.pi)
    x = np.arange(n) * 2 * np.pi / n
    F = o.xtok(1e-2 * np.cos(2 * x[:, None] + 10 * x[None, :]))
    assert abs(F[2, 10] - 5e-3) < 1e-17
    F[2, 10] = 0
    assert np.abs(F).max() < 1e-17


def test_parseval_and_round_trip():
    n = 64
    o = oracle.Oracle(nx=n, ny=n, Lx=2 * np.pi, Ly=2 * np.pi)
    rng = np.random.default_rng(1)
    F = rng.standard_normal(o.shape_k) + 1j * rng.standard_normal(o.shape_k)
    F[0, 0] = F[0, 0].real
    nkx = o.s["nkx"]
    for i in range(1, o.s["nkxh"]):           # Hermitian ky = 0 column
        F[nkx - i, 0] = np.conj(F[i, 0])
    f = o.ktox(F)
    G = o.xtok(f)
    np.testing.assert_allclose(G, F, rtol=0, atol=1e-12)
    w = np.ones(o.shape_k[1]) * 2.0
    w[0] = 1.0
    assert abs(np.mean(f ** 2) - np.sum(w[None, :] * np.abs(F) ** 2)) < 1e-9 * np.mean(f ** 2)


def test_sizes_closed_form():
    for n in list(range(4, 300)) + [1024, 2048, 4096, 8192]:
        s = oracle.sizes(n, n)
        nkx = (n - 1) // 3 * 2 + 1
        assert s["nkx"] == nkx and s["nky"] == (n - 1) // 3 + 1
        assert s["nkxh"] == (nkx + 1) // 2 and s["nkxh2"] == n - s["nkxh"]
        assert s["ncy"] == n // 2 + 1
        # retained |kx| < n/3 exactly (2/3 rule): the largest kept mode is (n-1)//3
        assert s["nkxh"] - 1 == (n - 1) // 3
    # published sizes at 2048^2 and 4096^2
    assert oracle.sizes(2048, 2048)["M"] == 932295 and oracle.sizes(2048, 2048)["I"] == 1398784
    assert oracle.sizes(4096, 4096)["M"] == 3730546
