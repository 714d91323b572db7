"""FP64 restatement of the learned spectral-convolution closure -- test
infrastructure only (tests and the CPU baseline may use it; the product path
never does).

The network has no counterpart in the reference, so its pins are mathematical:
the analysis basis against numpy's FFT (tests/test_closure_oracle.py), the
backward pass against central finite differences of the forward pass, and the
translation equivariance of a spectral layer.  Parameter layout, mode order and
the basis normalisation are those of include/shrt.h (shrt_closure_*):

    P[C][4], p[C]; per layer a[M][C][C], b[M][C][C], W[C][C], w[C]; Q[2][C], q[2]
    modes: ky = 0 with kx = 0..kmax, then each ky = 1..kmax with kx = -kmax..kmax
    basis column 2m = cos, 2m+1 = sin of mode m, orthonormal over the n*n grid
"""
import numpy as np

GELU_C = 0.7978845608028654


def modes(kmax):
    out = []
    for ky in range(kmax + 1):
        for kx in (range(0, kmax + 1) if ky == 0 else range(-kmax, kmax + 1)):
            out.append((kx, ky))
    return out


def basis(n, kmax):
    """E[n*n, 2M]: orthonormal cos/sin basis of the retained modes, points in
    the physical layout p = ix*n + iy."""
    ml = modes(kmax)
    ix, iy = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    ix, iy = ix.ravel(), iy.ravel()
    E = np.zeros((n * n, 2 * len(ml)))
    for m, (kx, ky) in enumerate(ml):
        th = 2 * np.pi * ((kx * ix + ky * iy) % n) / n
        if kx == 0 and ky == 0:
            E[:, 2 * m] = 1.0 / n
        else:
            s = np.sqrt(2.0) / n
            E[:, 2 * m] = s * np.cos(th)
            E[:, 2 * m + 1] = s * np.sin(th)
    return E


class Layout:
    def __init__(self, C, L, M):
        self.C, self.L, self.M = C, L, M
        o = 0
        self.P = o; o += 4 * C
        self.p = o; o += C
        self.a, self.b, self.W, self.w = [], [], [], []
        for _ in range(L):
            self.a.append(o); o += M * C * C
            self.b.append(o); o += M * C * C
            self.W.append(o); o += C * C
            self.w.append(o); o += C
        self.Q = o; o += 2 * C
        self.q = o; o += 2
        self.n = o

    def unpack(self, theta):
        C, M = self.C, self.M
        t = np.asarray(theta, dtype=np.float64)
        d = dict(P=t[self.P:self.P + 4 * C].reshape(C, 4), p=t[self.p:self.p + C],
                 Q=t[self.Q:self.Q + 2 * C].reshape(2, C), q=t[self.q:self.q + 2])
        d["a"] = [t[o:o + M * C * C].reshape(M, C, C) for o in self.a]
        d["b"] = [t[o:o + M * C * C].reshape(M, C, C) for o in self.b]
        d["W"] = [t[o:o + C * C].reshape(C, C) for o in self.W]
        d["w"] = [t[o:o + C] for o in self.w]
        return d


def bf16(x):
    """Round to BF16 (nearest even) through FP32, as the GPU's GEMMs round
    their operands; returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1 + np.tanh(GELU_C * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    th = np.tanh(GELU_C * (x + 0.044715 * x ** 3))
    return 0.5 * (1 + th) + 0.5 * x * (1 - th * th) * GELU_C * (1 + 3 * 0.044715 * x * x)


def factors(n, kmax):
    """The separable factors of the basis: Fy[n][2(kmax+1)] = cos, sin of
    2 pi ky iy/n (ky = 0..kmax); Fxc[n][2(2kmax+1)] = s cos, s sin and
    Fxs = -s sin, s cos of 2 pi kx ix/n (kx = -kmax..kmax), s = sqrt(2)/n.
    E's column pair of mode (kx, ky) is sum over the two stages, with the
    (0, 0) mode's cosine scaled by 1/sqrt(2)."""
    i = np.arange(n)
    ky = np.arange(kmax + 1)
    kx = np.arange(-kmax, kmax + 1)
    b = 2 * np.pi * ((np.outer(i, ky)) % n) / n
    a = 2 * np.pi * ((np.outer(i, kx)) % n) / n
    Fy = np.empty((n, 2 * (kmax + 1)))
    Fy[:, 0::2], Fy[:, 1::2] = np.cos(b), np.sin(b)
    s = np.sqrt(2.0) / n
    Fxc = np.empty((n, 2 * (2 * kmax + 1)))
    Fxs = np.empty_like(Fxc)
    Fxc[:, 0::2], Fxc[:, 1::2] = s * np.cos(a), s * np.sin(a)
    Fxs[:, 0::2], Fxs[:, 1::2] = -s * np.sin(a), s * np.cos(a)
    return Fy, Fxc, Fxs


class ClosureOracle:
    """Forward and hand-written backward pass in FP64.  With bf16=True every
    contraction rounds both operands to BF16 first, as the GPU's GEMMs do --
    including the two stages (along y, then along x) in which the GPU applies
    the basis -- so that the two agree to FP32 accumulation order rather than
    to BF16 rounding."""

    def __init__(self, n, kmax, width, layers, bf16=False):
        self.n, self.kmax, self.C, self.L = n, kmax, width, layers
        self.E = basis(n, kmax)
        self.M = self.E.shape[1] // 2
        self.lay = Layout(width, layers, self.M)
        self.bf = bf16
        self.q = globals()["bf16"] if bf16 else (lambda v: v)
        Fy, Fxc, Fxs = factors(n, kmax)
        self.Fy, self.Fxc, self.Fxs = self.q(Fy), self.q(Fxc), self.q(Fxs)

    def analysis(self, h):
        """X = h E for h [..., n*n]; in BF16 mode as the GPU's two GEMM stages."""
        if not self.bf:
            return h @ self.E
        q, n, K1 = self.q, self.n, self.kmax + 1
        T = q(h).reshape(h.shape[:-1] + (n, n)) @ self.Fy            # [..., ix, (ky, c)]
        Tc, Ts = q(T[..., 0::2]), q(T[..., 1::2])                     # [..., ix, ky]
        O = np.swapaxes(Tc, -1, -2) @ self.Fxc + np.swapaxes(Ts, -1, -2) @ self.Fxs   # [..., ky, J]
        X = O.reshape(h.shape[:-1] + (-1,))[..., 2 * self.kmax:].copy()
        X[..., 0] *= np.sqrt(0.5)
        return X

    def synthesis(self, Y):
        """Y E^T for coefficient rows Y [..., 2M]; in BF16 mode as the GPU's two stages."""
        if not self.bf:
            return Y @ self.E.T
        q, n, K1 = self.q, self.n, self.kmax + 1
        J = 2 * (2 * self.kmax + 1)
        Yf = np.zeros(Y.shape[:-1] + (K1 * J,))
        Yf[..., 2 * self.kmax:] = Y
        Yf[..., 2 * self.kmax] *= np.sqrt(0.5)
        Yr = q(Yf).reshape(Y.shape[:-1] + (K1, J))
        U = np.empty(Y.shape[:-1] + (n, 2 * K1))
        U[..., 0::2] = np.swapaxes(Yr @ self.Fxc.T, -1, -2)          # [..., ix, ky]
        U[..., 1::2] = np.swapaxes(Yr @ self.Fxs.T, -1, -2)
        z = q(U) @ self.Fy.T                                          # [..., ix, iy]
        return z.reshape(Y.shape[:-1] + (n * n,))

    def spectral(self, h, a, b):
        """K h for h [B][C][NP]: per mode Yc = Xc a^T - Xs b^T, Ys = Xc b^T + Xs a^T."""
        q = self.q
        X = self.analysis(h)                            # [B][C][2M]
        Xc, Xs = q(X[..., 0::2]), q(X[..., 1::2])       # [B][C][M]
        a, b = q(a), q(b)
        Yc = np.einsum("moi,bim->bom", a, Xc) - np.einsum("moi,bim->bom", b, Xs)
        Ys = np.einsum("moi,bim->bom", b, Xc) + np.einsum("moi,bim->bom", a, Xs)
        Y = np.empty_like(X)
        Y[..., 0::2], Y[..., 1::2] = Yc, Ys
        return self.synthesis(Y), X

    def forward(self, theta, x, keep=False):
        q = self.q
        d = self.lay.unpack(theta)
        x = np.asarray(x, dtype=np.float64)             # [B][4][NP]
        h = np.einsum("ci,bip->bcp", q(d["P"]), q(x)) + d["p"][None, :, None]
        hs, zs, Xs = [h], [], []
        for l in range(self.L):
            s, X = self.spectral(h, d["a"][l], d["b"][l])
            z = s + np.einsum("oi,bip->bop", q(d["W"][l]), q(h)) + d["w"][l][None, :, None]
            h = gelu(z)
            hs.append(h)
            zs.append(z)
            Xs.append(X)
        out = np.einsum("oc,bcp->bop", q(d["Q"]), q(h)) + d["q"][None, :, None]
        if keep:
            return out, (x, hs, zs, Xs, d)
        return out

    def loss_grad(self, theta, x, target):
        """MSE loss over all outputs and its gradient (flat, parameter layout)."""
        out, (x, hs, zs, Xs, d) = self.forward(theta, x, keep=True)
        cnt = out.size
        diff = out - np.asarray(target, dtype=np.float64)
        loss = float(np.sum(diff * diff) / cnt)
        dout = 2.0 * diff / cnt
        q = self.q
        lay, g = self.lay, np.zeros(self.lay.n)
        C, M = self.C, self.M
        g[lay.Q:lay.Q + 2 * C] = np.einsum("bop,bcp->oc", q(dout), q(hs[-1])).ravel()
        g[lay.q:lay.q + 2] = dout.sum(axis=(0, 2))
        dh = np.einsum("oc,bop->bcp", q(d["Q"]), q(dout))
        for l in reversed(range(self.L)):
            dz = dh * gelu_grad(zs[l])
            g[lay.w[l]:lay.w[l] + C] = dz.sum(axis=(0, 2))
            g[lay.W[l]:lay.W[l] + C * C] = np.einsum("bop,bip->oi", q(dz), q(hs[l])).ravel()
            dh = np.einsum("oi,bop->bip", q(d["W"][l]), q(dz))
            dY = self.analysis(dz)
            dYc, dYs = q(dY[..., 0::2]), q(dY[..., 1::2])
            X = Xs[l]
            Xc, Xsn = q(X[..., 0::2]), q(X[..., 1::2])
            a, b = q(d["a"][l]), q(d["b"][l])
            dXc = np.einsum("moi,bom->bim", a, dYc) + np.einsum("moi,bom->bim", b, dYs)
            dXs = -np.einsum("moi,bom->bim", b, dYc) + np.einsum("moi,bom->bim", a, dYs)
            ga = np.einsum("bom,bim->moi", dYc, Xc) + np.einsum("bom,bim->moi", dYs, Xsn)
            gb = -np.einsum("bom,bim->moi", dYc, Xsn) + np.einsum("bom,bim->moi", dYs, Xc)
            g[lay.a[l]:lay.a[l] + M * C * C] = ga.ravel()
            g[lay.b[l]:lay.b[l] + M * C * C] = gb.ravel()
            dX = np.empty_like(X)
            dX[..., 0::2], dX[..., 1::2] = dXc, dXs
            dh = dh + self.synthesis(dX)
        g[lay.P:lay.P + 4 * C] = np.einsum("bcp,bip->ci", q(dh), q(x)).ravel()
        g[lay.p:lay.p + C] = dh.sum(axis=(0, 2))
        return loss, g
