"""Pins the oracle's time stepping, shear bookkeeping, bracket and remap against what the
mathematics fixes: closed-form solutions (linear Rayleigh-Taylor growth, viscous decay of a
shearing wave), convergence order, conservation laws of the dealiased bracket, a closed-form
Jacobian of single modes, the exact FP64 remap step indices, and a brute-force remap by
physical wavenumber.  The explicit weights of the AB2/AB3 updates in BDF2 form are
pinned by the matrix exponential of the linear system with diffusion, and the sign of the
advection term by a one-step closed-form tendency of trigonometric fields."""
import numpy as np
import pytest
import scipy.linalg

from conftest import oracle

TWO_PI = 2 * np.pi


def mk(n=16, **kw):
    p = dict(Lx=TWO_PI, Ly=TWO_PI, noise=0, cfl_num=10.0, nu=0.0, kappa=0.0, sigma=0.0, g=1.0,
             rho0=1.0, rho0_prime=1.0, delt=1e-3)
    p.update(kw)
    return oracle.Oracle(nx=n, ny=n, **p)


@pytest.mark.parametrize("sigma,delt,steps,expect", [
    (0.5, 1e-3, 2000, [1999]),          # 128^2 small case: fires on the 2000th step
    (1.0, 2.5e-4, 8000, [3999, 7999]),  # 1024^2 case: fires at 4000 and 8000
    (1.5, 5e-5, 2000, []),              # 4096^2 case: never fires in 2000 steps
])
def test_remap_trigger_steps(sigma, delt, steps, expect):
    """The shift accumulates s -= sigma*kx*delt in FP64 without FMA; the remap fires
    only when s[1] lands within 1e-10 of -kx[1].  Independent of the grid size."""
    o = mk(16, sigma=sigma, delt=delt, linear=1)
    fired = []
    done = 0
    while done < steps:
        o.step(1)
        done += 1
        st = o.state()
        if st["remap_count"] > len(fired):
            fired.append(st["last_remap_step"])
    assert fired == expect


def test_linear_rayleigh_taylor_growth():
    """linear eq, sigma = nu = kappa = 0: rho(t) = eps*cosh(gamma t), gamma^2 = g rho0' kx^2/(rho0 k^2),
    omega(t) = -i (g/rho0) kx eps sinh(gamma t)/gamma; global error O(delt^2)."""
    kx, ky, eps, T = 1, 1, 1e-3, 1.0
    gamma = np.sqrt(kx * kx / (kx * kx + ky * ky))
    errs = []
    for dt in (2e-3, 1e-3):
        o = mk(16, linear=1, delt=dt)
        F = np.zeros(o.shape_k, complex)
        F[kx, ky] = eps
        o.set_spectral("rho", F)
        o.step(int(round(T / dt)))
        t = o.state()["time"]
        r = o.buffer("arho")[kx, ky]
        w = o.buffer("aomg")[kx, ky]
        r_ex = eps * np.cosh(gamma * t)
        w_ex = -1j * kx * eps * np.sinh(gamma * t) / gamma
        errs.append(max(abs(r - r_ex) / abs(r_ex), abs(w - w_ex) / abs(w_ex)))
    assert errs[1] < 1e-5
    assert 3.0 < errs[0] / errs[1] < 5.0   # second order


def test_shearing_wave_viscous_decay():
    """g = rho0' = 0, sigma != 0: rho_hat(t) = rho_hat(0) exp(-kappa * int (kx^2 + (ky - sigma kx t)^2) dt).
    The reference advances the shift after the diffusion (update_shear is the last call of
    time_advance), so the implicit term sees k_s(t_n): the scheme is first order here, with
    leading relative error kappa * delt * (k_s^2(T) - k_s^2(0)) in closed form."""
    kx, ky, sigma, kappa, T = 1, 3, 0.5, 0.05, 1.0
    errs = []
    for dt in (2e-3, 1e-3):
        o = mk(16, linear=1, g=0.0, rho0_prime=0.0, sigma=sigma, kappa=kappa, nu=kappa, delt=dt)
        F = np.zeros(o.shape_k, complex)
        F[kx, ky] = 1e-3
        o.set_spectral("rho", F)
        o.step(int(round(T / dt)))
        t = o.state()["time"]
        integral = kx * kx * t + (ky ** 2 * t - ky * sigma * kx * t ** 2 + (sigma * kx) ** 2 * t ** 3 / 3)
        ex = 1e-3 * np.exp(-kappa * integral)
        got = o.buffer("arho")[kx, ky]
        errs.append(abs(got - ex) / ex)
        lag = kappa * dt * abs((ky - sigma * kx * t) ** 2 - ky ** 2)
        assert abs(errs[-1] - lag) < 0.05 * lag
        assert abs(got) < ex            # k_s^2 falls here, so the lagged value over-damps
    assert 1.8 < errs[0] / errs[1] < 2.2


def test_diffusion_order():
    """Pure diffusion through Euler/BDF2 start-up then SBDF: second-order convergence."""
    errs = []
    for dt in (1e-2, 5e-3):
        o = mk(16, linear=1, g=0.0, rho0_prime=0.0, nu=0.1, kappa=0.1, delt=dt)
        F = np.zeros(o.shape_k, complex)
        F[2, 1] = 1.0
        o.set_spectral("rho", F)
        o.step(int(round(0.5 / dt)))
        t = o.state()["time"]
        errs.append(abs(o.buffer("arho")[2, 1] - np.exp(-0.1 * 5 * t)))
    assert 3.0 < errs[0] / errs[1] < 5.0


def test_invariants_of_the_dealiased_bracket():
    """nu = kappa = g = rho0' = sigma = 0: the truncated brackets conserve enstrophy,
    energy and density variance; only the time stepping breaks them (tiny drift)."""
    n = 32
    o = mk(n, g=0.0, rho0_prime=0.0, delt=1e-3)
    rng = np.random.default_rng(0)
    nkx, nky = o.shape_k
    W = np.zeros((nkx, nky), complex)
    R = np.zeros((nkx, nky), complex)
    for (i, j) in [(1, 2), (2, 1), (3, 3), (nkx - 2, 1), (1, 4), (nkx - 1, 2)]:
        W[i, j] = 0.1 * (rng.standard_normal() + 1j * rng.standard_normal())
        R[i, j] = 0.1 * (rng.standard_normal() + 1j * rng.standard_normal())
    o.set_spectral("omega", W)
    o.set_spectral("rho", R)
    d0 = o.diagnostics()
    o.step(200)
    d1 = o.diagnostics()
    for k in ("energy", "enstrophy", "rho_var"):
        assert abs(d1[k] - d0[k]) < 1e-6 * d0[k], k
    # the bracket does move energy between modes
    assert np.abs(o.buffer("aomg") - W).max() > 1e-4


def test_bracket_closed_form():
    """B(rho) = vx drho/dx + vy drho/dy with vx = dphi/dy, vy = -dphi/dx for
    phi = A cos(x + 2y), rho = C sin(3x + y): the Jacobian in closed form."""
    n = 32
    o = mk(n)
    nkx, nky = o.shape_k
    x = np.arange(n) * TWO_PI / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    A, Cc = 0.3, 0.2
    # omega = -lap(phi) = 5 A cos(x + 2y) -> omega_hat(1, 2) = 5A/2
    W = np.zeros((nkx, nky), complex)
    W[1, 2] = 5 * A / 2
    o.set_spectral("omega", W)
    o.check_cfl()
    Rh = o.xtok(Cc * np.sin(3 * X + Y))
    B = o.poisson_bracket(Rh)
    phi_y = -2 * A * np.sin(X + 2 * Y)
    phi_x = -A * np.sin(X + 2 * Y)
    rho_x = 3 * Cc * np.cos(3 * X + Y)
    rho_y = Cc * np.cos(3 * X + Y)
    np.testing.assert_allclose(B, phi_y * rho_x - phi_x * rho_y, atol=1e-14)
    # and the velocities themselves
    np.testing.assert_allclose(o.physical("vx"), phi_y, atol=1e-14)
    np.testing.assert_allclose(o.physical("vy"), -phi_x, atol=1e-14)


def test_bracket_closed_form_with_shift():
    """The same Jacobian in the sheared frame: with ky_shift s(kx) = kx*s1 the y
    derivatives take ky + s(kx) (ddy_shear) and phi = omega/(kx^2 + (ky + s)^2),
    while the grid values keep the stored ky.  phi = A cos(x + 2y) with shift s1,
    rho = C sin(3x + y) with shift 3*s1."""
    n = 32
    o = mk(n)
    nkx, nky = o.shape_k
    kx, _ = o.k_tables()
    s1 = 0.3
    o.set_shift(kx * s1)
    x = np.arange(n) * TWO_PI / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    A, Cc = 0.3, 0.2
    W = np.zeros((nkx, nky), complex)
    W[1, 2] = (1.0 + (2.0 + s1) ** 2) * A / 2
    o.set_spectral("omega", W)
    o.check_cfl()
    Rh = o.xtok(Cc * np.sin(3 * X + Y))
    B = o.poisson_bracket(Rh)
    vx = -(2.0 + s1) * A * np.sin(X + 2 * Y)
    vy = A * np.sin(X + 2 * Y)
    np.testing.assert_allclose(o.physical("vx"), vx, atol=1e-14)
    np.testing.assert_allclose(o.physical("vy"), vy, atol=1e-14)
    rho_x = 3 * Cc * np.cos(3 * X + Y)
    rho_y = (1.0 + 3 * s1) * Cc * np.cos(3 * X + Y)
    np.testing.assert_allclose(B, vx * rho_x + vy * rho_y, atol=1e-14)


def test_diagnostics_are_domain_means():
    """Energy, enstrophy and rho variance of the oracle (Hermitian-weighted sums
    over the stored half plane) equal the domain means of |grad phi|^2/2
    (sheared gradient), omega^2/2 and rho^2/2 of closed-form fields, including
    a ky = 0 mode (stored twice, weight 1) and a ky shift."""
    n = 32
    o = mk(n)
    nkx, nky = o.shape_k
    kx, _ = o.k_tables()
    s1 = -0.2
    o.set_shift(kx * s1)
    x = np.arange(n) * TWO_PI / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    A, Cc, D = 0.3, 0.5, 0.7
    ks2 = 1.0 + (2.0 + s1) ** 2
    W = np.zeros((nkx, nky), complex)
    W[1, 2] = ks2 * A / 2                      # phi = A cos(x + 2y), omega = ks2 * phi
    o.set_spectral("omega", W)
    o.set_spectral("rho", o.xtok(Cc * np.cos(2 * X) + D * np.sin(X + 3 * Y)))
    d = o.diagnostics()
    phi_x = -A * np.sin(X + 2 * Y)
    phi_y = -(2.0 + s1) * A * np.sin(X + 2 * Y)
    omega = ks2 * A * np.cos(X + 2 * Y)
    rho = Cc * np.cos(2 * X) + D * np.sin(X + 3 * Y)
    assert d["energy"] == pytest.approx(np.mean(0.5 * (phi_x ** 2 + phi_y ** 2)), rel=1e-13)
    assert d["enstrophy"] == pytest.approx(np.mean(0.5 * omega ** 2), rel=1e-13)
    assert d["rho_var"] == pytest.approx(np.mean(0.5 * rho ** 2), rel=1e-13)


def test_remap_matches_physical_wavenumber_definition():
    """After a remap each stored mode holds the old mode with the same physical
    wavenumber (kx, ky + s(kx)), or the conjugate of its Hermitian partner, or 0."""
    n = 32
    o = mk(n, sigma=1.0, delt=0.01, linear=1, g=0.0, rho0_prime=0.0)
    nkx, nky = o.shape_k
    nkxh = o.s["nkxh"]
    rng = np.random.default_rng(4)
    F = rng.standard_normal((nkx, nky)) + 1j * rng.standard_normal((nkx, nky))
    F[0, 0] = F[0, 0].real
    for i in range(1, nkxh):                       # a real field: Hermitian ky = 0 column
        F[nkx - i, 0] = np.conj(F[i, 0])
    # integer wavenumbers (Lx = Ly = 2 pi); sigma*t = 1 gives the old shift s = -kx
    kxi = np.where(np.arange(nkx) < nkxh, np.arange(nkx), np.arange(nkx) - nkx)
    s_old = -kxi
    jump = s_old.astype(np.int32)                  # round(s / kx[1]) with kx[1] = 1
    G = o.shearing_field(jump, F)
    # every old stored entry and its Hermitian partner, keyed by physical wavenumber
    table = {}
    for a in range(nkx):
        for b in range(nky):
            for key, val in (((kxi[a], b + s_old[a]), F[a, b]), ((-kxi[a], -(b + s_old[a])), np.conj(F[a, b]))):
                table.setdefault(key, []).append(val)
    for i in range(nkx):
        for j in range(nky):
            cands = table.get((kxi[i], j), [])     # new frame: s = 0
            want = cands[0] if cands else 0.0
            assert all(abs(c - want) < 1e-15 for c in cands)
            assert abs(G[i, j] - want) < 1e-15, (i, j)


def test_reference_noise_mask():
    """init_dis perturbs exactly the modes with |kx| <= 2 and 0 < |ky| <= 10, by at most rho_eps2."""
    n = 64
    o = oracle.Oracle(nx=n, ny=n, Lx=np.pi, Ly=np.pi, noise=1, rho_eps1=1e-2)
    o.initialize_reference()
    R = o.buffer("arho")
    kx, ky = o.k_tables()
    x = np.arange(n) * np.pi / n
    clean = o.xtok(1e-2 * np.cos(2 * x[:, None] + 10 * x[None, :]))
    d = R - clean
    mask = (np.abs(kx)[:, None] <= 2) & (np.abs(ky)[None, :] != 0) & (np.abs(ky)[None, :] <= 10)
    assert np.all(d[~mask] == 0)
    assert np.count_nonzero(d[mask]) == mask.sum()
    assert np.abs(d.real).max() <= 1e-3 and np.abs(d.imag).max() <= 1e-3


def test_cfl_halving_count():
    """delt halves while (g/rho0)*delt/dx > cfl: 0.1 with dx = 2pi/32, cfl 0.2 -> two halvings
    of that check (velocity checks are zero for omega = 0 and sigma = 0)."""
    o = mk(32, delt=0.1, cfl_num=0.2)
    o.step(1)
    st = o.state()
    dx = TWO_PI / 32
    k = 0
    dt = 0.1
    while dt / dx > 0.2:
        dt /= 2
        k += 1
    assert st["halvings"] == k == 2
    assert st["delt"] == dt


def _linear_exact(kx, ky, nu, kappa, g, rho0, rho0p, w0, r0, t):
    """Exact solution of the linear (linear eq = true, sigma = 0) system for one mode:
    d omega/dt = -(g/rho0) i kx rho - nu k^2 omega,  d rho/dt = rho0' i kx omega/k^2 - kappa k^2 rho
    (add_ddxrho, add_ddxphi with phi = omega/k^2, and the implicit diffusion), from the
    matrix exponential of its 2x2 generator."""
    k2 = kx * kx + ky * ky
    A = np.array([[-nu * k2, -(g / rho0) * 1j * kx],
                  [rho0p * 1j * kx / k2, -kappa * k2]], dtype=complex)
    return scipy.linalg.expm(A * t) @ np.array([w0, r0], dtype=complex)


@pytest.mark.parametrize("nu,kappa", [(0.1, 0.1), (0.05, 0.0), (0.0, 0.08), (0.03, 0.12)])
def test_linear_system_with_diffusion(nu, kappa):
    """Every scheme branch with nonzero tendencies: Euler forward/backward at step 0,
    AB2 and AB3 in BDF2 form with implicit BDF2 diffusion where the diffusivity is
    positive (time_integral.cu:633-646, 678-695, 747-767), plain AB2/AB3 where it is zero
    (:621-632, 664-677).  The buoyancy terms drive both fields, so the explicit weights act
    on nonzero tendencies; the global error must be second order against the matrix
    exponential.  A wrong weight (an inconsistent sum or a first-order start-up step)
    leaves an O(1) or O(delt) error and fails the order check."""
    kx, ky, eps, T = 1, 1, 1e-3, 1.0
    errs = []
    for dt in (2e-3, 1e-3):
        o = mk(16, linear=1, delt=dt, nu=nu, kappa=kappa)
        F = np.zeros(o.shape_k, complex)
        F[kx, ky] = eps
        o.set_spectral("rho", F)
        o.step(int(round(T / dt)))
        t = o.state()["time"]
        w_ex, r_ex = _linear_exact(kx, ky, nu, kappa, 1.0, 1.0, 1.0, 0.0, eps, t)
        w = o.buffer("aomg")[kx, ky]
        r = o.buffer("arho")[kx, ky]
        errs.append(max(abs(r - r_ex) / abs(r_ex), abs(w - w_ex) / abs(w_ex)))
        # the other modes stay exactly zero
        Wf = o.buffer("aomg").copy()
        Wf[kx, ky] = 0
        assert np.abs(Wf).max() == 0.0
    assert errs[1] < 2e-5, errs
    assert 3.3 < errs[0] / errs[1] < 4.7, errs   # second order


def test_linear_system_unequal_modes_and_damping():
    """The same closed form on two modes of different |k| at once, nu != kappa, with a
    nonzero initial vorticity: the per-mode kperp2 of the implicit division and the kx
    of the buoyancy terms are taken from the right rows."""
    modes = [(1, 2), (3, 1)]
    nu, kappa, T = 0.02, 0.05, 0.5
    errs = []
    for dt in (2e-3, 1e-3):
        o = mk(16, linear=1, delt=dt, nu=nu, kappa=kappa, g=2.0, rho0=1.5, rho0_prime=0.7)
        nkx = o.shape_k[0]
        W = np.zeros(o.shape_k, complex)
        R = np.zeros(o.shape_k, complex)
        for (a, b) in modes:
            W[a, b] = 1e-3 * (0.3 + 0.2j)
            R[a, b] = 1e-3 * (1.0 - 0.5j)
        # a kx < 0 row too: kx = -2 sits at row nkx - 2
        W[nkx - 2, 3] = 2e-4j
        R[nkx - 2, 3] = 5e-4
        o.set_spectral("omega", W)
        o.set_spectral("rho", R)
        o.step(int(round(T / dt)))
        t = o.state()["time"]
        e = 0.0
        for (row, kxv, b) in [(1, 1, 2), (3, 3, 1), (nkx - 2, -2, 3)]:
            w_ex, r_ex = _linear_exact(kxv, b, nu, kappa, 2.0, 1.5, 0.7, W[row, b], R[row, b], t)
            e = max(e, abs(o.buffer("aomg")[row, b] - w_ex) / abs(w_ex),
                    abs(o.buffer("arho")[row, b] - r_ex) / abs(r_ex))
        errs.append(e)
    assert errs[1] < 2e-5, errs
    assert 3.3 < errs[0] / errs[1] < 4.7, errs


def test_one_step_tendencies_closed_form():
    """The nonlinear tendencies of one step in closed form (time_integral.cu:488-617):
    d omega/dt = -(vx d omega/dx + vy d omega/dy) - (g/rho0) i kx rho_hat,
    d rho/dt   = -(vx d rho/dx + vy d rho/dy) + rho0' i kx phi_hat,
    vx = d phi/dy, vy = -d phi/dx (get_vector_shear), for
    phi = A cos(x + 2y) + D cos(3x + y) (omega = 5 A cos(x + 2y) + 10 D cos(3x + y)),
    rho = C sin(3x + y) + E cos(x - 2y).  The products stay inside the 2/3 truncation on
    32^2, so the spectra of the closed-form physical tendencies are exact.  Pins the sign
    of the advection term (negative(), fourier.cu:170-176) and of both buoyancy terms;
    the first step is Euler forward (nu = kappa = 0)."""
    n = 32
    dt = 1e-3
    o = mk(n, linear=0, delt=dt, g=1.3, rho0=1.1, rho0_prime=0.6)
    x = np.arange(n) * TWO_PI / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    A, D, C, E = 0.3, 0.1, 0.2, 0.15
    phi = A * np.cos(X + 2 * Y) + D * np.cos(3 * X + Y)
    om = 5 * A * np.cos(X + 2 * Y) + 10 * D * np.cos(3 * X + Y)
    rho = C * np.sin(3 * X + Y) + E * np.cos(X - 2 * Y)
    phi_x = -A * np.sin(X + 2 * Y) - 3 * D * np.sin(3 * X + Y)
    phi_y = -2 * A * np.sin(X + 2 * Y) - D * np.sin(3 * X + Y)
    om_x = -5 * A * np.sin(X + 2 * Y) - 30 * D * np.sin(3 * X + Y)
    om_y = -10 * A * np.sin(X + 2 * Y) - 10 * D * np.sin(3 * X + Y)
    rho_x = 3 * C * np.cos(3 * X + Y) - E * np.sin(X - 2 * Y)
    rho_y = C * np.cos(3 * X + Y) + 2 * E * np.sin(X - 2 * Y)
    vx, vy = phi_y, -phi_x
    W0 = o.xtok(om)
    R0 = o.xtok(rho)
    P0 = o.xtok(phi)
    o.set_spectral("omega", W0)
    o.set_spectral("rho", R0)
    np.testing.assert_allclose(o.buffer("aphi"), P0, atol=1e-15)
    o.step(1)
    kx, _ = o.k_tables()
    ikx = 1j * kx[:, None]
    dw_ex = o.xtok(-(vx * om_x + vy * om_y)) - (1.3 / 1.1) * ikx * R0
    dr_ex = o.xtok(-(vx * rho_x + vy * rho_y)) + 0.6 * ikx * P0
    # the bracket of omega is nonzero here (two modes of different |k|)
    assert np.abs(o.xtok(vx * om_x + vy * om_y)).max() > 1e-3
    np.testing.assert_allclose(o.buffer("domg"), dw_ex, atol=1e-14)
    np.testing.assert_allclose(o.buffer("drho"), dr_ex, atol=1e-14)
    np.testing.assert_allclose(o.buffer("aomg"), W0 + dt * dw_ex, atol=1e-15)
    np.testing.assert_allclose(o.buffer("arho"), R0 + dt * dr_ex, atol=1e-15)


def test_kdata_bef_snapshot_is_the_pre_remap_state():
    """The oracle keeps the fields update_shear hands to k_data_bef at a remap
    (shear.cu:236): remapped with the step's jump table they are the state after
    the remap, and before any remap the snapshot is empty."""
    n = 32
    o = mk(n, sigma=1.0, delt=0.01, linear=0, nu=1e-2, kappa=1e-2, cfl_num=5.0)
    o.set_spectral("rho", oracle.seeded_spectrum(n, n, TWO_PI, TWO_PI, 7, 1.0, 6.0, 1e-3, -1.0))
    o.step(99)
    assert np.abs(o.kdata_bef("rho")).max() == 0
    o.step(1)
    assert o.state()["last_remap_step"] == 99
    _, jump = o.shift()
    for f, ring in (("omega", "aomg"), ("phi", "aphi"), ("rho", "arho")):
        bef = o.kdata_bef(f)
        assert np.abs(bef).max() > 0
        np.testing.assert_array_equal(o.shearing_field(jump, bef), o.buffer(ring))
        assert not np.allclose(bef, o.buffer(ring))
