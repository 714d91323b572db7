"""Ensembles: independent instances of mixed sizes, precisions, shear rates and
seeds advanced by one captured graph per batch equal their individual runs
bit for bit (including a remap inside the run)."""
import numpy as np
import pytest

from conftest import shrt

pytestmark = pytest.mark.gpu

BASE = dict(Lx=2 * np.pi, Ly=2 * np.pi, nu=1e-3, kappa=1e-3, g=1.0, rho0=1.0, rho0_prime=1.0,
            cfl_num=0.35, noise=0, rho_eps1=1e-2)
SPECS = [  # nx, ny, precision, sigma, delt, seed
    (64, 64, shrt.FP32, 0.5, 1e-3, 1),
    (128, 128, shrt.FP32, 1.0, 1e-3, 2),
    (64, 128, shrt.FP64, 1.5, 1e-3, 3),
    (32, 32, shrt.FP64, 1.0, 1e-2, 4),     # remaps at step index 99
    (256, 256, shrt.FP32, 0.3, 5e-4, 5),
]


def make(nx, ny, prec, sigma, delt, seed, graph_steps=0, fuse=True):
    p = dict(BASE)
    p.update(sigma=sigma, delt=delt, cfl_num=5.0 if delt > 5e-3 else 0.35)
    g = shrt.Shrt(nx=nx, ny=ny, precision=prec, graph_steps=graph_steps, **p)
    g.set_fusion(fuse)
    g.initialize_seeded(seed=seed, kmin=2.0, kmax=8.0, amp=1e-3, expo=-1.0)
    return g


@pytest.mark.parametrize("fuse", [False, True])
def test_ensemble_equals_individual_runs(gpu, fuse):
    """Unfused: against single steps.  Fused: against each member's own
    captured batches of the same length (the same fused kernel sequence)."""
    print("ensemble of", len(SPECS), "fuse", fuse, flush=True)
    members = [make(*s, fuse=fuse) for s in SPECS]
    solo = [make(*s, graph_steps=4 if fuse else 0, fuse=fuse) for s in SPECS]
    ens = shrt.Ensemble(members, steps_per_graph=4)
    try:
        ens.step(102)           # 25 captured graphs, then 2 single steps per member
        for g in solo:
            g.time_advance(102)
        for spec, a, b in zip(SPECS, members, solo):
            for f in (shrt.OMEGA, shrt.RHO, shrt.PHI):
                assert np.array_equal(a.spectral(f).view(np.uint8), b.spectral(f).view(np.uint8)), (spec, f)
            assert a.state() == b.state(), spec
        assert members[3].state()["remap_count"] == 1
    finally:
        ens.close()
        for g in members + solo:
            g.close()


def test_ensemble_rejects_bad_arguments(gpu):
    g = make(*SPECS[0])
    try:
        with pytest.raises(shrt.ShrtError):
            shrt.Ensemble([g], steps_per_graph=3)
        with pytest.raises(shrt.ShrtError):
            shrt.Ensemble([g, g], steps_per_graph=4)
    finally:
        g.close()
