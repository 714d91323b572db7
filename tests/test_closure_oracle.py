"""Pins the FP64 restatement of the learned closure (oracle/closure_oracle.py)
by what the mathematics fixes: the analysis basis against numpy's FFT and its
orthonormality, the hand-written backward pass against central finite
differences of the forward pass, and the translation equivariance of the
network (a shifted input gives the shifted output)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import closure_oracle as co  # noqa: E402


@pytest.mark.parametrize("n,kmax", [(8, 2), (16, 5), (32, 7)])
def test_basis_is_fft_and_orthonormal(n, kmax):
    E = co.basis(n, kmax)
    ml = co.modes(kmax)
    f = np.random.default_rng(n).standard_normal((n, n))
    X = f.ravel() @ E
    F = np.fft.fft2(f)
    for m, (kx, ky) in enumerate(ml):
        s = 1.0 / n if (kx, ky) == (0, 0) else np.sqrt(2.0) / n
        assert abs(X[2 * m] - s * F[kx % n, ky].real) < 1e-12
        assert abs(X[2 * m + 1] + (0.0 if (kx, ky) == (0, 0) else s * F[kx % n, ky].imag)) < 1e-12
    G = E.T @ E
    expect = np.eye(E.shape[1])
    expect[1, 1] = 0.0       # the (0, 0) mode has no sine
    assert np.abs(G - expect).max() < 1e-12


def test_mode_count():
    assert len(co.modes(16)) == 17 + 16 * 33
    assert co.modes(1) == [(0, 0), (1, 0), (-1, 1), (0, 1), (1, 1)]


def _random_theta(o, rng, scale=0.3):
    return rng.standard_normal(o.lay.n) * scale


def test_backward_matches_finite_differences():
    o = co.ClosureOracle(n=8, kmax=2, width=3, layers=2)
    rng = np.random.default_rng(5)
    th = _random_theta(o, rng)
    x = rng.standard_normal((2, 4, 64))
    t = rng.standard_normal((2, 2, 64))
    loss, g = o.loss_grad(th, x, t)
    fd = np.zeros_like(th)
    h = 1e-6
    for i in range(th.size):
        e = np.zeros_like(th)
        e[i] = h
        lp, _ = o.loss_grad(th + e, x, t)
        lm, _ = o.loss_grad(th - e, x, t)
        fd[i] = (lp - lm) / (2 * h)
    assert np.abs(g - fd).max() < 1e-7 * max(1.0, np.abs(fd).max())
    # every parameter group receives a gradient (the imaginary weights of the
    # (0, 0) mode, which has no sine, excepted: they multiply zeros)
    lay = o.lay
    for off, cnt in [(lay.P, 12), (lay.a[0], 9), (lay.b[1] + 9, 9), (lay.W[1], 9), (lay.Q, 6), (lay.q, 2)]:
        assert np.abs(g[off:off + cnt]).max() > 0
    assert np.abs(g[lay.b[0]:lay.b[0] + 9]).max() == 0


def test_translation_equivariance():
    n = 16
    o = co.ClosureOracle(n=n, kmax=4, width=4, layers=2)
    rng = np.random.default_rng(9)
    th = _random_theta(o, rng)
    x = rng.standard_normal((1, 4, n, n))
    sx = np.roll(x, (3, -5), axis=(2, 3))
    y = o.forward(th, x.reshape(1, 4, n * n)).reshape(1, 2, n, n)
    ys = o.forward(th, sx.reshape(1, 4, n * n)).reshape(1, 2, n, n)
    assert np.abs(np.roll(y, (3, -5), axis=(2, 3)) - ys).max() < 1e-12


@pytest.mark.parametrize("n,kmax", [(16, 5), (32, 10), (64, 8)])
def test_separable_factors_reproduce_the_basis(n, kmax):
    """The two-stage form the GPU applies (a y stage of cos/sin of ky, then an x
    stage of cos/sin of kx, the (0, 0) cosine scaled by 1/sqrt(2)) equals the
    dense orthonormal basis: analysis h E and synthesis Y E^T."""
    E = co.basis(n, kmax)
    exact = co.ClosureOracle(n, kmax, 2, 1)
    rng = np.random.default_rng(n + kmax)
    h = rng.standard_normal((3, n * n))
    Fy, Fxc, Fxs = co.factors(n, kmax)
    T = h.reshape(3, n, n) @ Fy
    O = np.swapaxes(T[..., 0::2], -1, -2) @ Fxc + np.swapaxes(T[..., 1::2], -1, -2) @ Fxs
    X = O.reshape(3, -1)[:, 2 * kmax:].copy()
    X[:, 0] *= np.sqrt(0.5)
    np.testing.assert_allclose(X, h @ E, atol=1e-12)
    # the BF16 emulation of the GPU's stages stays within BF16 rounding of the exact transform
    q = co.ClosureOracle(n, kmax, 2, 1, bf16=True)
    assert np.linalg.norm(q.analysis(h) - h @ E) < 1e-2 * np.linalg.norm(h @ E)
    Y = rng.standard_normal((3, E.shape[1]))
    assert np.linalg.norm(q.synthesis(Y) - Y @ E.T) < 1e-2 * np.linalg.norm(Y @ E.T)
    np.testing.assert_allclose(exact.synthesis(Y), Y @ E.T, atol=1e-12)
