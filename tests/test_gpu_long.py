"""The plain method's GPU case at its full length, and the benchmark's 2048^2
workload over 1000 steps (about three and one minutes, most of it the FP64
oracle on the host).

1024^2, Lx = Ly = 2 pi, S = 1 (remaps near steps 4000 and 8000), nu = kappa
= 2e-4, g = 1, dt = 2.5e-4, 10000 steps, omega = 0 and a seeded random-phase
rho on 2 <= |k| <= 32 with amplitude 1e-4 |k|^-1.5; FP64 and FP32 contexts
against one FP64 oracle run, fields and the energy / enstrophy / rho-variance
series compared every 1000 steps.  Each row of the comparison is printed and
appended to gpurun_out/parity_long.txt as it is made."""
import os

import numpy as np
import pytest

from conftest import REPO, oracle, rel_l2, shrt

pytestmark = [pytest.mark.gpu]


def _row(text):
    print(text, flush=True)
    d = os.path.join(REPO, "gpurun_out")
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "parity_long.txt"), "a") as f:
        f.write(text + "\n")


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
    _row("1024^2 x 10000: step | FP64 omega rho | FP32 omega rho | energy enstrophy rho_var (FP64 rel, FP32 rel) | remaps")
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
        _row(" | ".join(row))
        for prec, g in gs.items():
            st = g.state()
            assert st["istep"] == (c + 1) * chunk and st["remap_count"] == o.state()["remap_count"]
    _row(f"worst field error FP64 {worst[8]:.2e}, FP32 {worst[4]:.2e}")
    assert worst[8] < 1e-10
    assert worst[4] < 1e-5
    for g in gs.values():
        g.close()
    o.close()


def test_headline_workload_1000_steps(gpu):
    """The benchmark's workload itself: 2048^2, sigma = 1, nu = kappa = 1e-4,
    dt = 1e-4, omega = 0 and a seeded random-phase rho on 2 <= |k| <= 32
    (amplitude 1e-4 |k|^-1.5); FP32 (the benchmark's precision) and FP64
    contexts against one FP64 oracle run over 1000 steps, every 250 steps.
    A third context runs FP32 with the state update in FP32 arithmetic (the
    reference's FP32 build, shrt_set_update_precision(0)); its error is
    reported beside the default FP64-evaluated, stochastically rounded update."""
    n, chunk, nchunk = 2048, 250, 4
    p = dict(Lx=2 * np.pi, Ly=2 * np.pi, nu=1e-4, kappa=1e-4, sigma=1.0, g=1.0, rho0=1.0, rho0_prime=1.0,
             delt=1e-4, cfl_num=0.35, noise=0, rho_eps1=1e-2)
    ic = dict(seed=20250101, kmin=2.0, kmax=32.0, amp=1e-4, expo=-1.5)
    o = oracle.Oracle(threads=16, nx=n, ny=n, **p)
    o.set_spectral("rho", oracle.seeded_spectrum(n, n, 2 * np.pi, 2 * np.pi, **ic))
    gs = {prec: shrt.Shrt(nx=n, ny=n, precision=prec, graph_steps=20, **p) for prec in (8, 4)}
    for g in gs.values():
        g.initialize_seeded(**ic)
    narrow = shrt.Shrt(nx=n, ny=n, precision=4, graph_steps=20, **p)
    narrow.set_update_precision(False)
    narrow.initialize_seeded(**ic)
    worst_narrow = 0.0
    _row("2048^2 x 1000: step | FP64 omega rho | FP32 omega rho | FP32-arith omega rho | energy enstrophy rho_var "
         "(FP64 rel, FP32 rel)")
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
        narrow.time_advance(chunk)
        en = [rel_l2(narrow.spectral(f), o.buffer(r)) for f, r in ((shrt.OMEGA, "aomg"), (shrt.RHO, "arho"))]
        worst_narrow = max(worst_narrow, *en)
        row.append(f"FP32-arith {en[0]:.2e} {en[1]:.2e}")
        dd = {prec: g.diagnostics() for prec, g in gs.items()}
        row.append(" ".join(f"{k} {do[k]:.6e} ({abs(dd[8][k] - do[k]) / do[k]:.1e}, {abs(dd[4][k] - do[k]) / do[k]:.1e})"
                            for k in ("energy", "enstrophy", "rho_var")))
        _row(" | ".join(row))
        for g in gs.values():
            assert g.state()["istep"] == (c + 1) * chunk
    _row(f"worst field error FP64 {worst[8]:.2e}, FP32 {worst[4]:.2e}, FP32 with the update in FP32 "
         f"arithmetic {worst_narrow:.2e}")
    assert worst[8] < 1e-10
    assert worst[4] < 1e-5
    assert worst_narrow < 1e-4
    for g in list(gs.values()) + [narrow]:
        g.close()
    o.close()
