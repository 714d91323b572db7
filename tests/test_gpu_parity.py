"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

Tolerances (relative L2 against the FP64 oracle):
  FP64 path: 1e-12 per transform / step, 1e-10 after many steps
  FP32 path: 2e-6 per transform, 1e-5 after many steps
Index work (remap, trigger step, halving count) is bit-exact.
"""
import numpy as np
import pytest

from conftest import oracle, rel_l2, random_spectrum, shrt

pytestmark = pytest.mark.gpu

BASE = dict(Lx=2 * np.pi, Ly=2 * np.pi, nu=1e-3, kappa=1e-3, sigma=0.5, g=1.0, rho0=1.0,
            rho0_prime=1.0, delt=1e-3, cfl_num=0.35, noise=0, rho_eps1=1e-2)
TOL1 = {8: 1e-12, 4: 2e-6}
TOLN = {8: 1e-10, 4: 1e-5}


def pair(n, prec, ny=None, graph_steps=0, **kw):
    p = dict(BASE)
    p.update(kw)
    ny = ny or n
    o = oracle.Oracle(nx=n, ny=ny, **p)
    g = shrt.Shrt(nx=n, ny=ny, precision=prec, graph_steps=graph_steps, **p)
    return o, g


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("nx,ny", [(32, 32), (64, 64), (128, 128), (64, 128), (256, 256), (512, 512),
                                   (1024, 1024)])
def test_xtok_ktox(gpu, prec, nx, ny):
    print(f"xtok/ktox nx={nx} ny={ny} prec={prec}", flush=True)
    o, g = pair(nx, prec, ny=ny)
    rng = np.random.default_rng(nx * 7 + ny)
    f = rng.standard_normal((nx, ny))
    assert rel_l2(g.xtok(f), o.xtok(f)) < TOL1[prec]
    S = random_spectrum(rng, *o.shape_k, decay=0.5, amp=1.0)
    assert rel_l2(g.ktox(S), o.ktox(S)) < TOL1[prec]


@pytest.mark.parametrize("prec", [8, 4])
def test_xtok_single_mode(gpu, prec):
    # cos(2x + 10y) has exactly two nonzero retained modes, each rho_eps1/2.
    n = 64
    o, g = pair(n, prec)
    x = np.arange(n) * 2 * np.pi / n
    f = 1e-2 * np.cos(2 * x[:, None] + 10 * x[None, :])
    F = g.xtok(f)
    assert abs(F[2, 10] - 5e-3) < 1e-8
    assert abs(F[o.s["nkx"] - 2, 10]) < 1e-8
    F[2, 10] = 0
    assert np.abs(F).max() < 1e-8


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("istep", [0, 1, 2, 7])
@pytest.mark.parametrize("nu,kappa", [(1e-3, 1e-3), (0.0, 1e-3), (1e-3, 0.0), (0.0, 0.0)])
def test_one_step_random_state(gpu, prec, istep, nu, kappa):
    """One time_advance from a random sheared state, every scheme branch: Euler,
    AB2/AB3 in BDF2 form with implicit diffusion where the diffusivity is positive,
    plain AB2/AB3 where it is zero (time_integral.cu:512-541, 575-605, 699-743)."""
    n = 64
    print(f"one step prec={prec} istep={istep} nu={nu} kappa={kappa}", flush=True)
    o, g = pair(n, prec, cfl_num=10.0, nu=nu, kappa=kappa)
    rng = np.random.default_rng(100 + istep)
    nkx, nky = o.shape_k
    g.set_time(istep, 0.1 * istep, 1e-3)
    o.set_time(istep, 0.1 * istep, 1e-3)
    kx, _ = o.k_tables()
    s = -0.5 * kx * 0.3
    s[0] = 0.0
    o.set_shift(s)
    g.set_shift(s)
    for ring in ("aomg", "domg", "arho", "drho"):
        for slot in (0, 1):
            amp = 1e-2 if ring[0] == "a" else 1e-3
            arr = random_spectrum(rng, nkx, nky, decay=2.0, amp=amp)
            if prec == 4:
                arr = arr.astype(np.complex64).astype(np.complex128)
            o.set_buffer(ring, slot, arr)
            g.set_buffer(ring, slot, arr)
    o.step(1)
    g.time_advance(1)
    for ring in ("aomg", "domg", "arho", "drho"):
        err = rel_l2(g.buffer(ring, 0), o.buffer(ring, 0))
        print(f"  {ring}0 rel err {err:.3e}")
        assert err < TOL1[prec] * 5, ring
    assert rel_l2(g.spectral(shrt.PHI), o.buffer("aphi")) < TOL1[prec] * 5
    so, sg = o.state(), g.state()
    assert sg["istep"] == so["istep"] == istep + 1
    assert sg["delt"] == so["delt"]
    sh_o, _ = o.shift()
    sh_g, _ = g.shift()
    assert np.array_equal(sh_o, sh_g)


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("n,steps", [(64, 200), (128, 100)])
def test_multistep_seeded(gpu, prec, n, steps):
    print(f"multistep n={n} prec={prec} steps={steps}", flush=True)
    o, g = pair(n, prec, sigma=0.5)
    ic = dict(seed=12345, kmin=2.0, kmax=8.0, amp=1e-3, expo=-1.0)
    o.set_spectral("rho", oracle.seeded_spectrum(n, n, 2 * np.pi, 2 * np.pi, **ic))
    g.initialize_seeded(**ic)
    r0 = g.spectral(shrt.RHO)
    assert rel_l2(r0, o.buffer("arho")) < (1e-16 if prec == 8 else 1e-7)
    o.step(steps)
    g.time_advance(steps)
    for f, ring in ((shrt.OMEGA, "aomg"), (shrt.RHO, "arho"), (shrt.PHI, "aphi")):
        err = rel_l2(g.spectral(f), o.buffer(ring))
        print(f"  field {f} rel err {err:.3e}")
        assert err < TOLN[prec]
    so, sg = o.state(), g.state()
    assert sg["istep"] == so["istep"]
    assert abs(sg["time"] - so["time"]) < 1e-12


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("nu,kappa", [(0.0, 1e-3), (1e-3, 0.0), (0.0, 0.0)])
def test_multistep_inviscid_branches(gpu, prec, nu, kappa):
    """Many steps with zero viscosity and/or diffusivity (the plain AB2/AB3 updates with
    no implicit division), single steps and fused captured batches, through a remap."""
    n = 64
    print(f"multistep inviscid prec={prec} nu={nu} kappa={kappa}", flush=True)
    prm = dict(BASE, nu=nu, kappa=kappa, sigma=1.0, delt=2e-3)
    ic = dict(seed=777, kmin=2.0, kmax=10.0, amp=2e-3, expo=-1.0)
    steps = 600   # sigma*t crosses 1 at step 500: one remap
    o = oracle.Oracle(nx=n, ny=n, **prm)
    o.set_spectral("rho", oracle.seeded_spectrum(n, n, 2 * np.pi, 2 * np.pi, **ic))
    o.step(steps)
    assert o.state()["remap_count"] == 1
    for gs in (0, 20):
        g = shrt.Shrt(nx=n, ny=n, precision=prec, graph_steps=gs, **prm)
        g.set_fusion(gs > 0)
        g.initialize_seeded(**ic)
        g.time_advance(steps)
        for f, ring in ((shrt.OMEGA, "aomg"), (shrt.RHO, "arho"), (shrt.PHI, "aphi")):
            err = rel_l2(g.spectral(f), o.buffer(ring))
            print(f"  graph_steps={gs} field {f} rel err {err:.3e}")
            assert err < TOLN[prec]
        sg, so = g.state(), o.state()
        assert sg["istep"] == so["istep"] and sg["remap_count"] == so["remap_count"] == 1
        assert sg["last_remap_step"] == so["last_remap_step"]
        g.close()


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("nx,ny,steps", [(64, 128, 60), (512, 512, 8), (1024, 1024, 4), (2048, 2048, 3)])
def test_multistep_sizes(gpu, prec, nx, ny, steps):
    """Non-square grids and the large transform lengths (one-wave teams at 512
    and 1024, the headline length 2048) against the oracle."""
    print(f"multistep nx={nx} ny={ny} prec={prec} steps={steps}", flush=True)
    o, g = pair(nx, prec, ny=ny, sigma=1.0)
    ic = dict(seed=4242, kmin=2.0, kmax=24.0, amp=1e-3, expo=-1.0)
    o.set_spectral("rho", oracle.seeded_spectrum(nx, ny, 2 * np.pi, 2 * np.pi, **ic))
    o.step(steps)
    # single steps, and captured batches of two steps (P3 fused with the next P1)
    for gs in (0, 2):
        print(f"  graph_steps={gs}", flush=True)
        g = shrt.Shrt(nx=nx, ny=ny, precision=prec, graph_steps=gs, **dict(BASE, sigma=1.0))
        g.set_fusion(gs > 0)
        g.initialize_seeded(**ic)
        g.time_advance(steps)
        for f, ring in ((shrt.OMEGA, "aomg"), (shrt.RHO, "arho"), (shrt.PHI, "aphi")):
            err = rel_l2(g.spectral(f), o.buffer(ring))
            print(f"  field {f} rel err {err:.3e}")
            assert err < TOLN[prec]
        assert g.state()["istep"] == o.state()["istep"]
        g.close()


@pytest.mark.parametrize("prec", [8, 4])
def test_small_case_full_length(gpu, prec):
    """The plain small case at its full length: 128^2, Lx = Ly = 2 pi, nu = kappa
    = 1e-3, g = 1, shear 0.5 (a remap whenever 0.5 t crosses an integer), dt =
    1e-3, 2000 steps (Euler and AB2 start-up, then AB3/BDF2), omega = 0 and a
    seeded random-phase rho on 2 <= |k| <= 8 with amplitude 1e-3 |k|^-1.  Fields
    within 1e-10 (FP64) and 1e-5 (FP32) relative L2 of the FP64 oracle after
    1000 and after 2000 steps (the remap falls on the last step); the shift
    bit-exact; energy, enstrophy and rho variance of the final state within the
    same tolerances.  Captured batches of 20 steps, as the driver runs them."""
    n = 128
    print(f"small case full length prec={prec}", flush=True)
    p = dict(BASE, sigma=0.5, nu=1e-3, kappa=1e-3, delt=1e-3)
    o = oracle.Oracle(threads=8, nx=n, ny=n, **p)
    g = shrt.Shrt(nx=n, ny=n, precision=prec, graph_steps=20, **p)
    ic = dict(seed=20240601, kmin=2.0, kmax=8.0, amp=1e-3, expo=-1.0)
    o.set_spectral("rho", oracle.seeded_spectrum(n, n, 2 * np.pi, 2 * np.pi, **ic))
    g.initialize_seeded(**ic)
    tol = TOLN[prec]
    for steps in (1000, 2000):
        o.step(1000)
        g.time_advance(1000)
        for f, ring in ((shrt.OMEGA, "aomg"), (shrt.RHO, "arho"), (shrt.PHI, "aphi")):
            err = rel_l2(g.spectral(f), o.buffer(ring))
            print(f"  step {steps} field {f} rel err {err:.3e}")
            assert err < tol
    steps = 2000
    so, sg = o.state(), g.state()
    assert sg["istep"] == so["istep"] == steps
    assert sg["remap_count"] == so["remap_count"] and sg["halvings"] == so["halvings"]
    sh_o, _ = o.shift()
    sh_g, _ = g.shift()
    assert np.array_equal(sh_o, sh_g)
    dg, do = g.diagnostics(), o.diagnostics()
    for k in ("energy", "enstrophy", "rho_var"):
        print(f"  {k} {dg[k]:.12e} vs {do[k]:.12e}")
        assert abs(dg[k] - do[k]) <= tol * abs(do[k])
    g.close()
    o.close()


def test_fp32_update_precision_forms(gpu):
    """The small case (128^2, sigma = 0.5, nu = kappa = 1e-3, dt = 1e-3) over 1000
    steps in FP32 with both forms of the state update against the FP64 oracle:
    evaluated in FP64 and rounded stochastically (the default) and in FP32
    arithmetic (the reference's FP32 build).  Both are reported; the default
    holds the 1e-5 bound."""
    n = 128
    p = dict(BASE, sigma=0.5, nu=1e-3, kappa=1e-3, delt=1e-3)
    ic = dict(seed=20240601, kmin=2.0, kmax=8.0, amp=1e-3, expo=-1.0)
    o = oracle.Oracle(threads=8, nx=n, ny=n, **p)
    o.set_spectral("rho", oracle.seeded_spectrum(n, n, 2 * np.pi, 2 * np.pi, **ic))
    o.step(1000)
    errs = {}
    for wide in (True, False):
        g = shrt.Shrt(nx=n, ny=n, precision=4, graph_steps=20, **p)
        g.set_update_precision(wide)
        g.initialize_seeded(**ic)
        g.time_advance(1000)
        errs[wide] = max(rel_l2(g.spectral(f), o.buffer(r)) for f, r in ((shrt.OMEGA, "aomg"), (shrt.RHO, "arho")))
        g.close()
    print(f"128^2 x 1000 FP32: FP64-evaluated update {errs[True]:.2e}, FP32 arithmetic {errs[False]:.2e}", flush=True)
    assert errs[True] < 1e-5
    assert errs[False] < 1e-4
    assert errs[False] > errs[True]


@pytest.mark.parametrize("prec", [4, 8])
def test_large_grid_4096(gpu, prec):
    """4096^2 for three steps from a seeded state on 4 <= |k| <= 64 against the
    FP64 oracle, single steps and a captured batch of two: the precision form
    (FFT, bracket and update in FP32; shift, time and the diagnostics' sums in
    FP64) and the FP64 context (one-team kernels, 67.6 KB of LDS per team)."""
    n, steps = 4096, 3
    print(f"4096^2 prec={prec} vs oracle", flush=True)
    p = dict(BASE, sigma=1.5, nu=5e-5, kappa=5e-5, delt=5e-5)
    o = oracle.Oracle(threads=16, nx=n, ny=n, **p)
    ic = dict(seed=4096, kmin=4.0, kmax=64.0, amp=1e-4, expo=0.0)
    o.set_spectral("rho", oracle.seeded_spectrum(n, n, 2 * np.pi, 2 * np.pi, **ic))
    o.step(steps)
    for gs in (0, 2):
        g = shrt.Shrt(nx=n, ny=n, precision=prec, graph_steps=gs, **p)
        g.initialize_seeded(**ic)
        g.time_advance(steps)
        for f, ring in ((shrt.OMEGA, "aomg"), (shrt.RHO, "arho")):
            err = rel_l2(g.spectral(f), o.buffer(ring))
            print(f"  graph_steps={gs} field {f} rel err {err:.3e}")
            assert err < TOLN[prec]
        assert g.state()["istep"] == steps
        g.close()
    o.close()


def test_config3_precision_form_4096_2000_steps(gpu):
    """The precision-form configuration at full length: 4096^2, sigma = 1.5,
    nu = kappa = 5e-5, dt = 5e-5, 2000 steps, rho seeded on 4 <= |k| <= 64 with
    amplitude 1e-4, omega = 0.  The FP32 context (FFT, bracket and update in
    FP32; kx(t) phase, remap indices and the energy/enstrophy reductions in
    FP64) against the FP64 GPU run of the same case: relative L2 of the fields
    and relative difference of the compensated energy / enstrophy / rho
    variance sums, every 500 steps; shift and step bookkeeping identical."""
    n, chunk = 4096, 500
    p = dict(BASE, sigma=1.5, nu=5e-5, kappa=5e-5, delt=5e-5)
    ic = dict(seed=4096, kmin=4.0, kmax=64.0, amp=1e-4, expo=0.0)
    gs = {prec: shrt.Shrt(nx=n, ny=n, precision=prec, graph_steps=20, **p) for prec in (8, 4)}
    for g in gs.values():
        g.initialize_seeded(**ic)
    worst = 0.0
    for c in range(4):
        for g in gs.values():
            g.time_advance(chunk)
        e = [rel_l2(gs[4].spectral(f), gs[8].spectral(f)) for f in (shrt.OMEGA, shrt.RHO, shrt.PHI)]
        d8, d4 = gs[8].diagnostics(), gs[4].diagnostics()
        dd = {k: abs(d4[k] - d8[k]) / d8[k] for k in ("energy", "enstrophy", "rho_var")}
        print(f"  step {(c + 1) * chunk}: FP32 vs FP64 GPU omega {e[0]:.2e} rho {e[1]:.2e} phi {e[2]:.2e}; "
              f"energy {dd['energy']:.1e} enstrophy {dd['enstrophy']:.1e} rho_var {dd['rho_var']:.1e}", flush=True)
        worst = max(worst, *e)
        assert max(dd.values()) < 1e-5
        s8, s4 = gs[8].state(), gs[4].state()
        assert s8 == s4 and s8["istep"] == (c + 1) * chunk and s8["halvings"] == 0
        assert np.array_equal(gs[8].shift()[0], gs[4].shift()[0])
    print(f"  worst FP32 vs FP64 GPU field error after 2000 steps: {worst:.2e}", flush=True)
    assert worst < 1e-5
    for g in gs.values():
        g.close()


@pytest.mark.parametrize("prec", [8, 4])
def test_reference_ic_and_steps(gpu, prec):
    n = 64
    o, g = pair(n, prec, noise=1, Lx=np.pi, Ly=np.pi, sigma=1.0, kappa=8e-5, cfl_num=0.5)
    o.initialize_reference()
    g.initialize()
    assert rel_l2(g.spectral(shrt.RHO), o.buffer("arho")) < TOL1[prec]
    o.step(30)
    g.time_advance(30)
    for f, ring in ((shrt.OMEGA, "aomg"), (shrt.RHO, "arho")):
        assert rel_l2(g.spectral(f), o.buffer(ring)) < TOLN[prec]


@pytest.mark.parametrize("prec", [8, 4])
def test_remap_trigger(gpu, prec):
    """sigma*delt = 0.01 with kx[1] = 1: the remap fires on step index 99."""
    n = 32
    print(f"remap trigger prec={prec}", flush=True)
    o, g = pair(n, prec, sigma=1.0, delt=0.01, cfl_num=5.0, nu=1e-2, kappa=1e-2)
    ic = dict(seed=7, kmin=1.0, kmax=6.0, amp=1e-3, expo=-1.0)
    o.set_spectral("rho", oracle.seeded_spectrum(n, n, 2 * np.pi, 2 * np.pi, **ic))
    g.initialize_seeded(**ic)
    o.step(100)
    g.time_advance(100)
    so, sg = o.state(), g.state()
    assert so["remap_count"] == 1 and so["last_remap_step"] == 99
    assert sg["remap_count"] == 1 and sg["last_remap_step"] == 99
    sh_o, j_o = o.shift()
    sh_g, j_g = g.shift()
    assert np.array_equal(sh_o, sh_g)
    assert np.array_equal(j_o, j_g)
    # right after the remap every live buffer is read through the remap view
    for ring in ("aomg", "arho", "domg", "drho"):
        for slot in (0, 1):
            err = rel_l2(g.buffer(ring, slot), o.buffer(ring, slot))
            assert err < TOLN[prec], (ring, slot, err)
    assert rel_l2(g.spectral(shrt.PHI), o.buffer("aphi")) < TOLN[prec]
    o.step(20)
    g.time_advance(20)
    for f, ring in ((shrt.OMEGA, "aomg"), (shrt.RHO, "arho")):
        assert rel_l2(g.spectral(f), o.buffer(ring)) < TOLN[prec]


def _remapped_pair(prec):
    """Oracle and GPU context right after a remap step (index 99), every live
    buffer of the GPU context still one frame behind (read through the view)."""
    o, g = pair(32, prec, sigma=1.0, delt=0.01, cfl_num=5.0, nu=1e-2, kappa=1e-2)
    ic = dict(seed=7, kmin=1.0, kmax=6.0, amp=1e-3, expo=-1.0)
    o.set_spectral("rho", oracle.seeded_spectrum(32, 32, 2 * np.pi, 2 * np.pi, **ic))
    g.initialize_seeded(**ic)
    o.step(100)
    g.time_advance(100)
    assert o.state()["last_remap_step"] == g.state()["last_remap_step"] == 99
    return o, g


def _all_buffers_match(o, g, tol):
    for ring in ("aomg", "arho", "domg", "drho"):
        for slot in (0, 1, 2):
            err = rel_l2(g.buffer(ring, slot), o.buffer(ring, slot))
            assert err < tol, (ring, slot, err)
    assert rel_l2(g.spectral(shrt.PHI), o.buffer("aphi")) < tol


@pytest.mark.parametrize("prec", [8, 4])
def test_spectral_before_remap(gpu, prec):
    """Right after a remap step the stored buffers are the fields update_shear hands
    to k_data_bef (shear.cu:236): the oracle's state with the remap of that step
    undone, i.e. its buffers before shearing_field; the plain getters return the
    remapped ones (k_data_aft)."""
    print(f"spectral before remap prec={prec}", flush=True)
    o, g = pair(32, prec, sigma=1.0, delt=0.01, cfl_num=5.0, nu=1e-2, kappa=1e-2)
    ic = dict(seed=7, kmin=1.0, kmax=6.0, amp=1e-3, expo=-1.0)
    o.set_spectral("rho", oracle.seeded_spectrum(32, 32, 2 * np.pi, 2 * np.pi, **ic))
    g.initialize_seeded(**ic)
    o.step(99)
    g.time_advance(99)
    for f in (shrt.OMEGA, shrt.RHO, shrt.PHI):     # no remap pending: both getters agree
        assert np.array_equal(g.spectral_unremapped(f), g.spectral(f))
    o.step(1)
    g.time_advance(1)
    assert o.state()["last_remap_step"] == g.state()["last_remap_step"] == 99
    # the stored (unremapped) buffers, remapped with the step's jump table by the
    # oracle's shearing_field, are the oracle's state after the remap
    _, jump = o.shift()
    for f, ring in ((shrt.OMEGA, "aomg"), (shrt.RHO, "arho"), (shrt.PHI, "aphi")):
        raw = g.spectral_unremapped(f).astype(np.complex128)
        assert rel_l2(raw, o.kdata_bef({shrt.OMEGA: "omega", shrt.RHO: "rho", shrt.PHI: "phi"}[f])) < TOLN[prec], f
        fwd = o.shearing_field(jump, raw)
        assert rel_l2(fwd, o.buffer(ring)) < TOLN[prec], f
        assert rel_l2(g.spectral(f), o.buffer(ring)) < TOLN[prec], f
        assert not np.allclose(raw, g.spectral(f))


@pytest.mark.parametrize("prec", [8, 4])
def test_spectral_modes_filter(gpu, prec):
    """shrt_get_spectral_modes equals cutting the coarse grid's retained modes out
    of the full spectrum on the host (rows kx >= 0 and kx < 0 of the coarse grid,
    its first nky modes), bit for bit, also while a remap is pending."""
    import importlib.util, os
    from conftest import REPO
    spec = importlib.util.spec_from_file_location("closure_pipe", os.path.join(REPO, "shearing-rt_amd", "closure.py"))
    cl = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(cl)
    o, g = _remapped_pair(prec)   # right after a remap: the view applies
    for nc in (8, 16, 32):
        filt = cl.Filter(shrt.sizes(32, 32), shrt.sizes(nc, nc))
        for f in (shrt.OMEGA, shrt.RHO, shrt.PHI):
            got = g.spectral_modes(f, nc)
            assert np.array_equal(got.view(np.uint8), filt(g.spectral(f)).view(np.uint8)), (nc, f)
    with pytest.raises(RuntimeError):
        g.spectral_modes(shrt.OMEGA, 64)


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("new_istep", [100, 101, 105, 2])
def test_set_time_keeps_state(gpu, prec, new_istep):
    """shrt_set_time right after a remap, with a step count in each class mod 3 (the
    ring slots are addressed by it): every named buffer keeps its contents and its
    pending remap view, and the following steps match the oracle, whose set_time only
    changes the counters.  A new step size beyond |sigma|*delt < 1/3 is refused."""
    print(f"set_time prec={prec} istep={new_istep}", flush=True)
    o, g = _remapped_pair(prec)
    t = o.state()["time"]
    o.set_time(new_istep, t, 0.005)
    g.set_time(new_istep, t, 0.005)
    _all_buffers_match(o, g, TOLN[prec])
    st = g.state()
    assert st["istep"] == new_istep and st["delt"] == 0.005
    o.step(7)
    g.time_advance(7)
    for f, ring in ((shrt.OMEGA, "aomg"), (shrt.RHO, "arho"), (shrt.PHI, "aphi")):
        assert rel_l2(g.spectral(f), o.buffer(ring)) < TOLN[prec]
    with pytest.raises(RuntimeError):
        g.set_time(new_istep, t, 0.34)


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("what", ["omega", "rho", "ring_aomg", "ring_drho"])
def test_set_one_field_after_remap(gpu, prec, what):
    """Right after a remap both fields of the state slot are pending; setting one of
    them (or one buffer of a ring slot) must leave the other's remap in place.  The
    next steps match the oracle given the same assignment."""
    print(f"set one field after remap prec={prec} {what}", flush=True)
    o, g = _remapped_pair(prec)
    rng = np.random.default_rng(5)
    arr = random_spectrum(rng, *o.shape_k, decay=2.0, amp=1e-3)
    if prec == 4:
        arr = arr.astype(np.complex64).astype(np.complex128)
    if what in ("omega", "rho"):
        o.set_spectral(what, arr)
        g.set_spectral(shrt.OMEGA if what == "omega" else shrt.RHO, arr)
    else:
        ring = what[5:]
        o.set_buffer(ring, 1, arr)
        g.set_buffer(ring, 1, arr)
    _all_buffers_match(o, g, TOLN[prec])
    o.step(3)
    g.time_advance(3)
    for f, ring in ((shrt.OMEGA, "aomg"), (shrt.RHO, "arho")):
        assert rel_l2(g.spectral(f), o.buffer(ring)) < TOLN[prec]


@pytest.mark.parametrize("prec,n", [(8, 1024), (4, 1024), (8, 2048), (4, 2048), (4, 4096)])
def test_remap_large_grid(gpu, prec, n):
    """A remap on the large grids, where nkx - 1 exceeds 1024 from 2048^2 up (the size
    at which the reference's per-block reset in shearing_ky races, shear.cu:275-294;
    both sides implement its race-free reading).  The shift starts three steps short of
    sigma*t = 1, so the remap fires on step index 2 with jumps up to nkx - 1 (rows
    beyond nky empty, the conjugate fold for kx < 0): trigger step, shift and jump
    table bit-exact, fields within tolerance before and after, single steps and fused
    captured batches."""
    print(f"large-grid remap n={n} prec={prec}", flush=True)
    sigma = 1.0
    dt = 1e-4 if n < 4096 else 5e-5    # max|vx + sigma*y|*dt/dx stays below the CFL number
    p = dict(BASE, sigma=sigma, nu=1e-4, kappa=1e-4, delt=dt)
    o = oracle.Oracle(threads=16, nx=n, ny=n, **p)
    kx, _ = o.k_tables()
    s0 = -(kx * (1.0 - 3 * sigma * dt))
    s0[0] = 0.0
    ic = dict(seed=31 + n, kmin=2.0, kmax=32.0, amp=1e-4, expo=-1.5)
    o.set_spectral("rho", oracle.seeded_spectrum(n, n, 2 * np.pi, 2 * np.pi, **ic))
    o.set_shift(s0)
    o.step(3)
    so = o.state()
    assert so["halvings"] == 0
    assert so["remap_count"] == 1 and so["last_remap_step"] == 2
    sh_o, j_o = o.shift()
    assert np.abs(j_o).max() == o.s["nkxh"] - 1
    ref3 = {r: o.buffer(r) for r in ("aomg", "arho", "aphi")}
    o.step(3)
    for gs in (0, 2):
        g = shrt.Shrt(nx=n, ny=n, precision=prec, graph_steps=gs, **p)
        g.set_fusion(gs > 0)
        g.initialize_seeded(**ic)
        g.set_shift(s0)
        g.time_advance(3)
        sg = g.state()
        assert sg["remap_count"] == 1 and sg["last_remap_step"] == 2
        sh_g, j_g = g.shift()
        assert np.array_equal(sh_o, sh_g) and np.array_equal(j_o, j_g)
        for f, ring in ((shrt.OMEGA, "aomg"), (shrt.RHO, "arho"), (shrt.PHI, "aphi")):
            err = rel_l2(g.spectral(f), ref3[ring])
            print(f"  graph_steps={gs} after the remap step, field {f} rel err {err:.3e}", flush=True)
            assert err < TOLN[prec]
        g.time_advance(3)
        for f, ring in ((shrt.OMEGA, "aomg"), (shrt.RHO, "arho")):
            err = rel_l2(g.spectral(f), o.buffer(ring))
            print(f"  graph_steps={gs} three steps later, field {f} rel err {err:.3e}", flush=True)
            assert err < TOLN[prec]
        g.close()
    o.close()


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("nx,ny,linear", [(32, 32, 0), (64, 128, 0), (128, 128, 0), (32, 32, 1), (2048, 2048, 0)])
def test_fused_batches_with_remaps(gpu, prec, nx, ny, linear):
    """Captured batches run P3 of step n and P1 of step n+1 as one kernel.  With
    sigma*delt*kx[1] = 0.2 a remap fires every fifth step (indices 4, 9, 14, 19):
    at every position of a four-step batch, so the fused kernel forms the remap
    view of the rows it has just updated (and the last step of a batch hands a
    pending remap to the next batch's P1).  Oracle parity, index work
    bit-exact, and the unfused batches and single steps agree to rounding;
    also at the headline size 2048^2, where FP32 contexts run fused by
    default."""
    print(f"fused batches nx={nx} ny={ny} linear={linear} prec={prec}", flush=True)
    p = dict(sigma=20.0, delt=0.01, cfl_num=1e3, nu=1e-2, kappa=1e-2, linear=linear)
    o = oracle.Oracle(nx=nx, ny=ny, threads=16 if nx >= 1024 else 1, **dict(BASE, **p))
    ic = dict(seed=17, kmin=1.0, kmax=8.0, amp=1e-3, expo=-1.0)
    o.set_spectral("rho", oracle.seeded_spectrum(nx, ny, 2 * np.pi, 2 * np.pi, **ic))
    steps = 22
    o.step(steps)
    so = o.state()
    assert so["remap_count"] == 4 and so["last_remap_step"] == 19
    outs = {}
    for mode in ("fused", "unfused", "single", "p1pair"):
        g = shrt.Shrt(nx=nx, ny=ny, precision=prec, graph_steps=0 if mode == "single" else 4, **dict(BASE, **p))
        g.set_fusion(mode == "fused")
        g.set_variants(2 if mode == "p1pair" else 0, 0, 0)
        g.initialize_seeded(**ic)
        g.time_advance(steps)
        sg = g.state()
        assert sg["remap_count"] == 4 and sg["last_remap_step"] == 19, mode
        assert sg["istep"] == so["istep"] and sg["halvings"] == so["halvings"] == 0
        sh_o, j_o = o.shift()
        sh_g, j_g = g.shift()
        assert np.array_equal(sh_o, sh_g) and np.array_equal(j_o, j_g), mode
        outs[mode] = [g.spectral(f) for f in (shrt.OMEGA, shrt.RHO, shrt.PHI)]
        for got, ring in zip(outs[mode], ("aomg", "arho", "aphi")):
            err = rel_l2(got, o.buffer(ring))
            print(f"  {mode} {ring} rel err {err:.3e}", flush=True)
            assert err < TOLN[prec], (mode, ring, err)
        g.close()
    for mode in ("unfused", "single", "p1pair"):
        for a, b in zip(outs["fused"], outs[mode]):
            same = np.array_equal(a.view(np.uint8), b.view(np.uint8))
            print(f"  fused vs {mode}: {'bit-identical' if same else f'rel {rel_l2(a, b):.2e}'}")
            assert same or rel_l2(a, b) < TOL1[prec]


@pytest.mark.parametrize("prec", [8, 4])
def test_shearing_field_bitexact(gpu, prec):
    n = 64
    o, g = pair(n, prec)
    nkx, nky = o.shape_k
    nkxh = o.s["nkxh"]
    rng = np.random.default_rng(3)
    dt = np.complex128 if prec == 8 else np.complex64
    F = (rng.standard_normal((nkx, nky)) + 1j * rng.standard_normal((nkx, nky))).astype(dt)
    # the jump table of a remap at sigma*t = 1, and a random one
    ikx = np.arange(nkx)
    kxi = np.where(ikx < nkxh, ikx, ikx - nkx)
    tables = [-kxi, rng.integers(-nky, nky, size=nkx)]
    tables[1][0] = 0
    for jmp in tables:
        jmp = jmp.astype(np.int32)
        ref = o.shearing_field(jmp, F.astype(np.complex128)).astype(dt)
        out = g.shearing_field(jmp, F)
        assert np.array_equal(out.view(np.uint8), ref.view(np.uint8))


@pytest.mark.parametrize("prec", [8, 4])
def test_cfl_halving(gpu, prec):
    """g/rho0 and velocity checks halve delt on the device exactly like check_cfl."""
    n = 32
    o, g = pair(n, prec, delt=0.1, cfl_num=0.2)
    ic = dict(seed=11, kmin=1.0, kmax=5.0, amp=1e-2, expo=-1.0)
    o.set_spectral("rho", oracle.seeded_spectrum(n, n, 2 * np.pi, 2 * np.pi, **ic))
    g.initialize_seeded(**ic)
    o.step(5)
    g.time_advance(5)
    so, sg = o.state(), g.state()
    assert so["halvings"] > 0
    assert sg["halvings"] == so["halvings"]
    assert sg["delt"] == so["delt"]
    recs = g.halvings()
    assert len(recs) >= 1
    assert sum(r["count"] for r in recs) == sg["halvings"]
    for r in recs:
        # the per-condition counts replay the printed halvings of check_cfl
        assert sum(r["counts"]) == r["count"]
        assert r["delt_before"] / 2.0 ** r["count"] == r["delt"]
        assert r["values"][2] == 1.0 and r["values"][3] == 1.0   # g/rho0, rho0_prime
    assert rel_l2(g.spectral(shrt.RHO), o.buffer("arho")) < TOLN[prec]


@pytest.mark.parametrize("prec", [8, 4])
def test_linear_mode(gpu, prec):
    n = 64
    o, g = pair(n, prec, linear=1)
    ic = dict(seed=5, kmin=2.0, kmax=10.0, amp=1e-3, expo=-1.0)
    o.set_spectral("rho", oracle.seeded_spectrum(n, n, 2 * np.pi, 2 * np.pi, **ic))
    g.initialize_seeded(**ic)
    o.step(50)
    g.time_advance(50)
    for f, ring in ((shrt.OMEGA, "aomg"), (shrt.RHO, "arho")):
        assert rel_l2(g.spectral(f), o.buffer(ring)) < TOLN[prec]


@pytest.mark.parametrize("prec", [8, 4])
def test_graph_replay_matches_eager(gpu, prec):
    """Unfused captured batches replay the single-step launches bit for bit."""
    n = 64
    ic = dict(seed=99, kmin=2.0, kmax=8.0, amp=1e-3, expo=-1.0)
    outs = []
    for gs in (0, 4):
        g = shrt.Shrt(nx=n, ny=n, precision=prec, graph_steps=gs, **BASE)
        g.set_fusion(False)
        g.initialize_seeded(**ic)
        g.time_advance(9)
        g.time_advance(13)
        outs.append((g.spectral(shrt.OMEGA), g.spectral(shrt.RHO), g.spectral(shrt.PHI), g.state()))
        g.close()
    for a, b in zip(outs[0][:3], outs[1][:3]):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    assert outs[0][3] == outs[1][3]


@pytest.mark.parametrize("prec", [8, 4])
def test_diagnostics_and_physical(gpu, prec):
    n = 64
    o, g = pair(n, prec)
    ic = dict(seed=21, kmin=2.0, kmax=8.0, amp=1e-3, expo=-1.0)
    o.set_spectral("rho", oracle.seeded_spectrum(n, n, 2 * np.pi, 2 * np.pi, **ic))
    g.initialize_seeded(**ic)
    o.step(40)
    g.time_advance(40)
    do, dg = o.diagnostics(), g.diagnostics()
    for k in ("energy", "enstrophy", "rho_var"):
        assert abs(dg[k] - do[k]) <= TOLN[prec] * abs(do[k]), k
    for pf, ring in ((shrt.P_OMEGA, "aomg"), (shrt.P_RHO, "arho"), (shrt.P_PHI, "aphi")):
        assert rel_l2(g.physical(pf), o.ktox(o.buffer(ring))) < TOLN[prec]
    assert rel_l2(g.physical(shrt.P_VX), o.physical("vx")) < TOLN[prec]
    assert rel_l2(g.physical(shrt.P_VY), o.physical("vy")) < TOLN[prec]


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("nx,ny", [(32, 32), (64, 64), (64, 128), (256, 256)])
def test_ktox_shear(gpu, prec, nx, ny):
    """Lab-frame output transform with a random shift against the oracle's
    direct sheared sum."""
    print(f"ktox_shear nx={nx} ny={ny} prec={prec}", flush=True)
    o, g = pair(nx, prec, ny=ny, sigma=1.0)
    nkx, nky = o.shape_k
    kx, _ = o.k_tables()
    rng = np.random.default_rng(nx * 3 + ny)
    s = rng.uniform(-0.9, 0.9, nkx) * kx[1]
    s[0] = 0.0
    o.set_shift(s)
    g.set_shift(s)
    S = random_spectrum(rng, nkx, nky, decay=0.5, amp=1.0)
    assert rel_l2(g.ktox_shear(S), o.ktox_shear(S)) < TOL1[prec] * 5


@pytest.mark.parametrize("prec", [8, 4])
def test_lab_frame_output(gpu, prec):
    """What output_fields writes when sigma != 0: omega, phi, rho in the lab
    frame after a run, including right after a remap (pending view)."""
    n = 32
    print(f"lab frame output prec={prec}", flush=True)
    o, g = pair(n, prec, sigma=1.0, delt=0.01, cfl_num=5.0, nu=1e-2, kappa=1e-2)
    ic = dict(seed=17, kmin=1.0, kmax=6.0, amp=1e-3, expo=-1.0)
    o.set_spectral("rho", oracle.seeded_spectrum(n, n, 2 * np.pi, 2 * np.pi, **ic))
    g.initialize_seeded(**ic)
    for steps in (37, 63):   # the second run ends on the remap step (index 99)
        o.step(steps)
        g.time_advance(steps)
        for pf, ring in ((shrt.P_OMEGA, "aomg"), (shrt.P_PHI, "aphi"), (shrt.P_RHO, "arho")):
            err = rel_l2(g.physical(pf, lab_frame=1), o.ktox_shear(o.buffer(ring)))
            assert err < TOLN[prec], (steps, pf, err)
    assert g.state()["remap_count"] == 1
    with pytest.raises(shrt.ShrtError):
        g.physical(shrt.P_VX, lab_frame=1)


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("nx,ny", [(32, 32), (64, 128)])
def test_bracket_output(gpu, prec, nx, ny):
    """The advection term that output_pbk writes: the bracket of the current
    state as the next step computes it (velocities of phi = omega/kperp2 with
    the current shift), before and right after a remap, against the oracle's
    check_cfl velocities and poisson_bracket."""
    print(f"bracket output nx={nx} ny={ny} prec={prec}", flush=True)
    o, g = pair(nx, prec, ny=ny, sigma=1.0, delt=0.01, cfl_num=5.0, nu=1e-2, kappa=1e-2)
    ic = dict(seed=19, kmin=1.0, kmax=6.0, amp=1e-1, expo=-1.0)
    o.set_spectral("rho", oracle.seeded_spectrum(nx, ny, 2 * np.pi, 2 * np.pi, **ic))
    g.initialize_seeded(**ic)
    for steps in (5, 32, 63):   # the last run ends on the remap step (index 99)
        o.step(steps)
        g.time_advance(steps)
        o.check_cfl()           # only sets aphi, vx, vy from the current omega
        ref = o.poisson_bracket(o.buffer("aomg"))
        got = g.physical(shrt.P_BRACKET)
        err = rel_l2(got, ref)
        assert err < TOLN[prec], (steps, err)
    assert g.state()["remap_count"] == 1
    with pytest.raises(shrt.ShrtError):
        g.physical(shrt.P_BRACKET, lab_frame=1)
    g.close()
    lin = shrt.Shrt(nx=32, ny=32, precision=prec, graph_steps=0, **dict(BASE, linear=1))
    with pytest.raises(shrt.ShrtError):
        lin.physical(shrt.P_BRACKET)
    lin.close()


@pytest.mark.parametrize("prec", [8, 4])
def test_kernel_variants(gpu, prec):
    """Every combination of the pass variants matches the oracle and the others
    (to rounding: the variants are compiled in different kernels, so the
    compiler may contract products differently); autotune picks valid ones."""
    n, steps = 128, 40
    print(f"variants prec={prec}", flush=True)
    ic = dict(seed=31, kmin=2.0, kmax=10.0, amp=1e-3, expo=-1.0)
    o = oracle.Oracle(nx=n, ny=n, **dict(BASE, sigma=1.0))
    o.set_spectral("rho", oracle.seeded_spectrum(n, n, 2 * np.pi, 2 * np.pi, **ic))
    o.step(steps)
    ref = o.buffer("arho")
    outs = {}
    for v in [(a, b, c) for a in (0, 1, 2) for b in (0, 1) for c in (0, 1, 2)]:
        g = shrt.Shrt(nx=n, ny=n, precision=prec, graph_steps=4, **dict(BASE, sigma=1.0))
        g.set_variants(*v)
        g.initialize_seeded(**ic)
        g.time_advance(steps)
        outs[v] = g.spectral(shrt.RHO)
        assert rel_l2(outs[v], ref) < TOLN[prec], v
        g.close()
    for v, r in outs.items():
        assert rel_l2(r, outs[(0, 0, 0)]) < TOLN[prec], v
    # one row of a pair per P3 workgroup: the same arithmetic per mode as the pair form
    for a in (0, 1, 2):
        for b in (0, 1):
            assert np.array_equal(outs[(a, b, 2)].view(np.uint8), outs[(a, b, 0)].view(np.uint8)), (a, b)
    g = shrt.Shrt(nx=n, ny=n, precision=prec, graph_steps=4, **dict(BASE, sigma=1.0))
    t = g.autotune(reps=5)
    assert t["p1"] in (0, 1, 2) and t["p2"] in (0, 1) and t["p3"] in (0, 1, 2) and t["fuse"] in (0, 1)
    assert all(m > 0 for m in t["ms"]) and all(m > 0 for m in t["step_ms"])
    assert t["fuse"] == (t["step_ms"][1] < t["step_ms"][0])
    g.initialize_seeded(**ic)     # the tuned context steps (fused or not) and matches the oracle
    g.time_advance(steps)
    assert rel_l2(g.spectral(shrt.RHO), ref) < TOLN[prec]
    g.close()


@pytest.mark.parametrize("n,prec", [(128, 8), (128, 4), (2048, 4)])
def test_fused_kernel_time_matches_a_fused_batch(gpu, n, prec):
    """shrt_fused_kernel_time launches nsteps + 1 steps exactly as a fused
    captured batch of that many steps does (at 2048^2 FP32 with the fused
    kernel's arguments in device memory): the state after it is the batch's,
    bit for bit, and the timing is positive."""
    print(f"fused kernel timing n={n} prec={prec}", flush=True)
    ic = dict(seed=17, kmin=2.0, kmax=8.0, amp=1e-3, expo=-1.0)
    g = shrt.Shrt(nx=n, ny=n, precision=prec, graph_steps=6, **dict(BASE, sigma=1.0))
    g.set_fusion(True)
    g.initialize_seeded(**ic)
    g.time_advance(6)
    ref = (g.spectral(shrt.OMEGA), g.spectral(shrt.RHO), g.spectral(shrt.PHI), g.state())
    g.close()
    g = shrt.Shrt(nx=n, ny=n, precision=prec, graph_steps=0, **dict(BASE, sigma=1.0))
    g.initialize_seeded(**ic)
    ms = g.fused_kernel_time(5)
    out = (g.spectral(shrt.OMEGA), g.spectral(shrt.RHO), g.spectral(shrt.PHI), g.state())
    g.close()
    assert ms > 0
    for a, b in zip(out[:3], ref[:3]):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    assert out[3] == ref[3] and out[3]["istep"] == 6


def test_copy_benches(gpu):
    """The bandwidth references of the bench: both copy kernels time a copy and
    reject sizes that are not whole 16-byte words."""
    import ctypes
    for fn in (shrt.lib().shrt_copy_bench, shrt.lib().shrt_stream_copy_bench):
        ms = ctypes.c_double(-1.0)
        assert fn(1 << 22, 3, ctypes.byref(ms)) == 0
        assert 0.0 < ms.value < 100.0
        assert fn(1000 * 16 + 8, 1, ctypes.byref(ms)) != 0
        assert fn(0, 1, ctypes.byref(ms)) != 0
