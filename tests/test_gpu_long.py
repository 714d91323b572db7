"""The plain method's GPU case at its full length (run on request: SHRT_LONG=1;
about three minutes, most of it the FP64 oracle on the host).

1024^2, Lx = Ly = 2 pi, S = 1 (remaps near steps 4000 and 8000), nu = kappa
= 2e-4, g = 1, dt = 2.5e-4, 10000 steps, omega = 0 and a seeded random-phase
rho on 2 <= |k| <= 32 with amplitude 1e-4 |k|^-1.5; FP64 and FP32 contexts
against one FP64 oracle run, fields and the energy / enstrophy / rho-variance
series compared every 1000 steps."""
import os

import numpy as np
import pytest

from conftest import oracle, rel_l2, shrt

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.environ.get("SHRT_LONG"), reason="long run: set SHRT_LONG=1")]


def test_gpu_case_10000_steps(gpu):
    n, chunk, nchunk = 1024, 1000, 10
    p = dict(Lx=2 * np.pi, Ly=2 * np.pi, nu=2e-4, kappa=2e-4, sigma=1.0, g=1.0, rho0=1.0, rho0_prime=1.0,
             delt=2.5e-4, cfl_num=0.35, noise=0, rho_eps1=1e-2)
    ic = dict(seed=1024, kmin=2.0, kmax=32.0, amp=1e-4, expo=-1.5)
    o = oracle.Oracle(threads=16, nx=n, ny=n, **p)
    o.set_spectral("rho", oracle.seeded_spectrum(n, n, 2 * np.pi, 2 * np.pi, **ic))
    gs = {prec: shrt.Shrt(nx=n, ny=n, precision=prec, graph_steps=20, **p) for prec in (8, 4)}
    for g in gs.values():
        g.initialize_seeded(**ic)
    print("step | FP64 omega rho | FP32 omega rho | energy enstrophy rho_var (FP64 rel, FP32 rel) | remaps")
    worst = {8: 0.0, 4: 0.0}
    for c in range(nchunk):
        o.step(chunk)
        do = o.diagnostics()
        row = [f"{(c + 1) * chunk:5d}"]
        for prec, g in gs.items():
            g.time_advance(chunk)
            e = [rel_l2(g.spectral(f), o.buffer(r)) for f, r in ((shrt.OMEGA, "aomg"), (shrt.RHO, "arho"))]
            worst[prec] = max(worst[prec], *e)
            row.append(f"{e[0]:.2e} {e[1]:.2e}")
        dd = {prec: g.diagnostics() for prec, g in gs.items()}
        row.append(" ".join(f"{k} {do[k]:.6e} ({abs(dd[8][k] - do[k]) / do[k]:.1e}, {abs(dd[4][k] - do[k]) / do[k]:.1e})"
                            for k in ("energy", "enstrophy", "rho_var")))
        row.append(str(o.state()["remap_count"]))
        print(" | ".join(row), flush=True)
        for prec, g in gs.items():
            st = g.state()
            assert st["istep"] == (c + 1) * chunk and st["remap_count"] == o.state()["remap_count"]
    print(f"worst field error FP64 {worst[8]:.2e}, FP32 {worst[4]:.2e}")
    assert worst[8] < 1e-10
    assert worst[4] < 1e-5
    for g in gs.values():
        g.close()
    o.close()


def test_headline_workload_1000_steps(gpu):
    """The benchmark's workload itself: 2048^2, sigma = 1, nu = kappa = 1e-4,
    dt = 1e-4, omega = 0 and a seeded random-phase rho on 2 <= |k| <= 32
    (amplitude 1e-4 |k|^-1.5); FP32 (the benchmark's precision) and FP64
    contexts against one FP64 oracle run over 1000 steps, every 250 steps."""
    n, chunk, nchunk = 2048, 250, 4
    p = dict(Lx=2 * np.pi, Ly=2 * np.pi, nu=1e-4, kappa=1e-4, sigma=1.0, g=1.0, rho0=1.0, rho0_prime=1.0,
             delt=1e-4, cfl_num=0.35, noise=0, rho_eps1=1e-2)
    ic = dict(seed=20250101, kmin=2.0, kmax=32.0, amp=1e-4, expo=-1.5)
    o = oracle.Oracle(threads=16, nx=n, ny=n, **p)
    o.set_spectral("rho", oracle.seeded_spectrum(n, n, 2 * np.pi, 2 * np.pi, **ic))
    gs = {prec: shrt.Shrt(nx=n, ny=n, precision=prec, graph_steps=20, **p) for prec in (8, 4)}
    for g in gs.values():
        g.initialize_seeded(**ic)
    print("step | FP64 omega rho | FP32 omega rho | energy enstrophy rho_var (FP64 rel, FP32 rel)")
    worst = {8: 0.0, 4: 0.0}
    for c in range(nchunk):
        o.step(chunk)
        do = o.diagnostics()
        row = [f"{(c + 1) * chunk:5d}"]
        for prec, g in gs.items():
            g.time_advance(chunk)
            e = [rel_l2(g.spectral(f), o.buffer(r)) for f, r in ((shrt.OMEGA, "aomg"), (shrt.RHO, "arho"))]
            worst[prec] = max(worst[prec], *e)
            row.append(f"{e[0]:.2e} {e[1]:.2e}")
        dd = {prec: g.diagnostics() for prec, g in gs.items()}
        row.append(" ".join(f"{k} {do[k]:.6e} ({abs(dd[8][k] - do[k]) / do[k]:.1e}, {abs(dd[4][k] - do[k]) / do[k]:.1e})"
                            for k in ("energy", "enstrophy", "rho_var")))
        print(" | ".join(row), flush=True)
        for g in gs.values():
            assert g.state()["istep"] == (c + 1) * chunk
    print(f"worst field error FP64 {worst[8]:.2e}, FP32 {worst[4]:.2e}")
    assert worst[8] < 1e-10
    assert worst[4] < 1e-5
    for g in gs.values():
        g.close()
    o.close()
