"""Pins the oracle's lab-frame output transform ktox_shear (shear.cu:467-527,
fft.cu:208-250) against what the mathematics fixes: with no shift it is the
ordinary inverse transform of a real field's spectrum, and a single stored mode
(kx, ky) with shift s becomes w * Re(A exp(i(kx x + (ky + s) y))) on the grid
x = ix*dx, y = iy*dy, with weight w = 1 for ky = 0 and 2 for ky > 0."""
import numpy as np
import pytest

from conftest import oracle

TWO_PI = 2 * np.pi


def mk(nx, ny, Lx=TWO_PI, Ly=TWO_PI):
    p = dict(Lx=Lx, Ly=Ly, noise=0, cfl_num=10.0, nu=0.0, kappa=0.0, sigma=1.0, g=1.0,
             rho0=1.0, rho0_prime=1.0, delt=1e-3)
    return oracle.Oracle(nx=nx, ny=ny, **p)


@pytest.mark.parametrize("nx,ny", [(16, 16), (32, 16), (24, 48)])
def test_unsheared_is_ktox(nx, ny):
    o = mk(nx, ny)
    o.set_shift(np.zeros(o.shape_k[0]))
    f = np.random.default_rng(nx + ny).standard_normal((nx, ny))
    F = o.xtok(f)
    assert np.allclose(o.ktox_shear(F), o.ktox(F), rtol=0, atol=1e-12 * np.abs(f).max())


@pytest.mark.parametrize("ikx,iky", [(3, 5), (0, 4), (2, 0), (-3, 2), (-1, 0)])
def test_single_sheared_mode(ikx, iky):
    nx, ny = 32, 32
    Lx, Ly = TWO_PI, 2.0
    o = mk(nx, ny, Lx, Ly)
    nkx, nky = o.shape_k
    kx, ky = o.k_tables()
    rng = np.random.default_rng(7)
    s = rng.uniform(-0.9, 0.9, nkx) * kx[1]
    s[0] = 0.0
    o.set_shift(s)
    row = ikx % nkx
    F = np.zeros((nkx, nky), dtype=np.complex128)
    A = 0.3 - 0.7j
    F[row, iky] = A
    x = np.arange(nx) * Lx / nx
    y = np.arange(ny) * Ly / ny
    w = 1.0 if iky == 0 else 2.0
    theta = kx[row] * x[:, None] + (ky[iky] + s[row]) * y[None, :]
    expect = w * np.real(A * np.exp(1j * theta))
    got = o.ktox_shear(F)
    assert np.abs(got - expect).max() < 1e-12


def test_shift_is_a_row_phase():
    """Shifting one row by s multiplies its contribution by exp(i s y) before
    the real part: two rows with opposite shifts stay independent."""
    o = mk(16, 16)
    nkx, nky = o.shape_k
    rng = np.random.default_rng(3)
    F = rng.standard_normal((nkx, nky)) + 1j * rng.standard_normal((nkx, nky))
    s = np.zeros(nkx)
    o.set_shift(s)
    base = o.ktox_shear(F)
    s[2] = 0.37
    o.set_shift(s)
    shifted = o.ktox_shear(F)
    G = np.zeros_like(F)
    G[2] = F[2]
    o.set_shift(np.zeros(nkx))
    row0 = o.ktox_shear(G)
    s2 = np.zeros(nkx)
    s2[2] = 0.37
    o.set_shift(s2)
    row1 = o.ktox_shear(G)
    assert np.allclose(shifted - base, row1 - row0, atol=1e-12)
