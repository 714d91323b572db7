"""The learned closure's network on the GPU (BF16 MFMA GEMMs, FP32
accumulation) against its FP64 restatement: forward outputs and every
parameter gradient, and Adam training that recovers a teacher network of the
same architecture.

Two references: the oracle with every contraction's operands rounded to BF16
as the GPU rounds them (agreement to FP32 accumulation order: relative L2
< 2e-3 for outputs and gradients, per parameter group < 1e-2), and the exact
FP64 oracle (agreement to BF16 rounding: outputs < 2e-2; gradients, which sum
many rounded products with cancellation and compound the rounding layer by
layer, to cosine > 0.9 -- measured 0.998 with two layers, 0.977 with three)."""
import os
import sys

import numpy as np
import pytest

from conftest import shrt

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import closure_oracle as co  # noqa: E402

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def inputs(rng, B, n):
    """Smooth random fields (a few low modes) plus a constant channel."""
    x = np.zeros((B, 4, n * n))
    ix, iy = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    for b in range(B):
        for c in range(3):
            f = np.zeros((n, n))
            for _ in range(6):
                kx, ky = rng.integers(-4, 5, 2)
                f += rng.standard_normal() * np.cos(2 * np.pi * (kx * ix + ky * iy) / n + rng.uniform(0, 2 * np.pi))
            x[b, c] = f.ravel()
        x[b, 3] = rng.uniform(0.5, 1.5)
    return x


@pytest.mark.parametrize("n,kmax,width,layers,B", [(16, 4, 8, 2, 2), (32, 6, 16, 3, 3), (64, 8, 32, 4, 2)])
def test_forward_and_gradients_match_oracle(gpu, n, kmax, width, layers, B):
    print(f"closure parity n={n} kmax={kmax} width={width} layers={layers} batch={B}", flush=True)
    net = shrt.Closure(n, kmax, width, layers, max_batch=4, seed=3)
    o = co.ClosureOracle(n, kmax, width, layers)
    ob = co.ClosureOracle(n, kmax, width, layers, bf16=True)
    assert net.nparams == o.lay.n and net.nmodes == o.M
    rng = np.random.default_rng(n)
    th = net.params().astype(np.float64)
    x = inputs(rng, B, n)
    y_gpu = net.forward(x)
    y_ref, y_bf = o.forward(th, x), ob.forward(th, x)
    print(f"  forward rel {rel(y_gpu, y_bf):.2e} (bf16 oracle) {rel(y_gpu, y_ref):.2e} (fp64 oracle)")
    assert rel(y_gpu, y_bf) < 2e-3
    assert rel(y_gpu, y_ref) < 2e-2
    t = rng.standard_normal((B, 2, n * n)) * np.std(y_ref) + y_ref
    loss_gpu, g_gpu = net.gradients(x, t)
    loss_ref, g_ref = o.loss_grad(th, x, t)
    loss_bf, g_bf = ob.loss_grad(th, x, t)
    cos = float(np.dot(g_gpu, g_ref) / (np.linalg.norm(g_gpu) * np.linalg.norm(g_ref)))
    print(f"  loss {loss_gpu:.6e} vs {loss_bf:.6e} (bf16) {loss_ref:.6e} (fp64); grad rel {rel(g_gpu, g_bf):.2e} "
          f"(bf16) {rel(g_gpu, g_ref):.2e} (fp64) cos {cos:.6f}")
    assert abs(loss_gpu - loss_bf) < 2e-3 * loss_bf
    assert rel(g_gpu, g_bf) < 2e-3
    assert cos > 0.9
    lay = o.lay
    for name, off, cnt in [("P", lay.P, 4 * width), ("p", lay.p, width), ("a0", lay.a[0], o.M * width * width),
                           ("b_last", lay.b[-1], o.M * width * width), ("W0", lay.W[0], width * width),
                           ("w_last", lay.w[-1], width), ("Q", lay.Q, 2 * width), ("q", lay.q, 2)]:
        r = rel(g_gpu[off:off + cnt], g_bf[off:off + cnt])
        print(f"    {name}: {r:.2e}")
        assert r < 1e-2, (name, r)
    net.close()


def test_training_recovers_a_teacher(gpu):
    """Targets from a teacher network of the same architecture: Adam on the
    device must cut the loss by well over an order of magnitude."""
    n, kmax, width, layers = 32, 6, 16, 2
    print("closure teacher/student", flush=True)
    rng = np.random.default_rng(11)
    teacher = co.ClosureOracle(n, kmax, width, layers)
    student = shrt.Closure(n, kmax, width, layers, max_batch=8, seed=21)
    th_t = shrt.Closure(n, kmax, width, layers, max_batch=1, seed=99)
    theta = th_t.params().astype(np.float64)
    th_t.close()
    x = inputs(rng, 48, n)
    y = teacher.forward(theta, x)
    student.set_data(x, y)
    losses = student.train(400, 8, 3e-3, seed=5)
    first, last = float(np.mean(losses[:10])), float(np.mean(losses[-10:]))
    print(f"  loss {first:.3e} -> {last:.3e}")
    assert np.all(np.isfinite(losses))
    assert last < first / 20
    # held-out inputs: the student generalises to the teacher's map
    xv = inputs(np.random.default_rng(12), 4, n)
    err = rel(student.forward(xv), teacher.forward(theta, xv))
    base = rel(np.zeros_like(teacher.forward(theta, xv)), teacher.forward(theta, xv))
    print(f"  held-out rel error {err:.3e} (zero predictor {base:.3e})")
    assert err < 0.3 * base
    student.close()


def test_closure_rejects_bad_shapes(gpu):
    with pytest.raises(shrt.ShrtError):
        shrt.Closure(24, 4)          # not a power of two
    with pytest.raises(shrt.ShrtError):
        shrt.Closure(16, 8)          # kmax must stay below n/2


def test_closure_beats_the_plain_coarse_solver(gpu):
    """The whole pipeline at small size: stably stratified sheared turbulence,
    128^2 fine runs filtered to 32^2 generate the data (48 trajectories x 40
    snapshots), the network trains on the device with early stopping on two
    validation trajectories, and on two held-out seeds the corrected 32^2 run
    must track the filtered 128^2 enstrophy spectrum better than the plain
    32^2 run (measured: about 0.2x its error, close to the plain 64^2 run)."""
    import importlib.util
    here = os.path.dirname(os.path.abspath(__file__))
    spec = importlib.util.spec_from_file_location("closure_pipe", os.path.join(here, "..", "shearing-rt_amd",
                                                                               "closure.py"))
    cl = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(cl)
    print("closure pipeline 128^2 -> 32^2", flush=True)
    cfg = cl.Config(rho0_prime=-1.0, amp=0.5, kmax_ic=8.0, spinup=500, nu=5e-4, snaps=40, width=16, kmax_net=8)
    p = cl.Pipeline(shrt, cfg)
    x, y = p.generate(range(100, 148))
    val = p.generate(range(800, 802))
    losses = p.fit(x, y, iters=1000, batch=16, lr=5e-4, val=val)
    assert np.all(np.isfinite(losses))
    print(f"  validation R2 curve {[(a, round(b, 3)) for a, b in p.val_curve]}")
    res = p.evaluate([900, 901], 500)
    for r in res:
        print(f"  seed {r['seed']}: enstrophy spectrum error plain {r['plain']['enstrophy_spectrum']:.3f} "
              f"closure {r['closure']['enstrophy_spectrum']:.3f} plain 64^2 {r['mid']['enstrophy_spectrum']:.3f}; "
              f"rho plain {r['plain']['rho']:.3f} closure {r['closure']['rho']:.3f}")
        assert r["skipped"] == 0 and r["skipped_device"] == 0
        # the device-resident correction makes the host path's arithmetic
        assert r["device_vs_host"] < 1e-6, r["device_vs_host"]
    plain = np.mean([r["plain"]["enstrophy_spectrum"] for r in res])
    closure = np.mean([r["closure"]["enstrophy_spectrum"] for r in res])
    assert closure < 0.6 * plain
    assert np.mean([r["closure"]["rho"] for r in res]) < np.mean([r["plain"]["rho"] for r in res])
    p.net.close()


def _pipeline_module():
    import importlib.util
    here = os.path.dirname(os.path.abspath(__file__))
    spec = importlib.util.spec_from_file_location("closure_pipe", os.path.join(here, "..", "shearing-rt_amd",
                                                                               "closure.py"))
    cl = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(cl)
    return cl


@pytest.mark.parametrize("prec,kcut", [(4, 0), (8, 0), (4, 5)])
def test_device_correction_matches_host(prec, kcut):
    """shrt_closure_correct against the same correction made through the host
    API (ktox of the stored fields, forward, xtok, set_spectral): the same
    state afterwards, bit for bit, also right after a remap (pending view) and
    over the steps that follow; the fallback (tiny guard, non-finite state)
    leaves the state unchanged.  With kcut the corrections keep |kx|, ky <= kcut
    on both paths."""
    cl = _pipeline_module()
    n = 32
    print(f"device closure correction prec={prec}", flush=True)
    cfg = cl.Config(n_coarse=n, n_fine=64, n_mid=64, delt=1e-2, interval=10, cfl=5.0, width=8, kmax_net=6,
                    layers=2, rho0_prime=-1.0, kcut=kcut)
    p = cl.Pipeline(shrt, cfg)
    p.net = shrt.Closure(n, 6, 8, 2, max_batch=1, seed=5)
    p.scale_in = np.array([0.7, 0.3, 0.05])
    p.scale_out = np.array([2.0, 0.5])
    sigma, dt = 1.0, cfg.interval * cfg.delt

    def make():
        c = shrt.Shrt(nx=n, ny=n, precision=prec, graph_steps=0, **cfg.params(sigma))
        c.initialize_seeded(seed=3, kmin=1.0, kmax=6.0, amp=0.3, expo=-1.0)
        return c

    def same(a, b):
        for f in (shrt.OMEGA, shrt.RHO, shrt.PHI):
            x, y = a.spectral(f), b.spectral(f)
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), f

    a, b = make(), make()
    for steps in (37, 63):   # the second run ends on the remap step (index 99)
        a.time_advance(steps)
        b.time_advance(steps)
        corr = p.correction(a, sigma)
        om, rh = a.spectral(shrt.OMEGA), a.spectral(shrt.RHO)
        a.set_spectral(shrt.OMEGA, om + dt * corr[0])
        a.set_spectral(shrt.RHO, rh + dt * corr[1])
        assert p.net.correct(b, p.scale_in, sigma, p.scale_out, dt, 10.0, kcut)
        same(a, b)
        assert np.linalg.norm(a.spectral(shrt.OMEGA) - om) > 0
    assert b.state()["remap_count"] == 1
    a.time_advance(15)
    b.time_advance(15)
    same(a, b)
    # fallback: a guard of zero rejects any correction
    before = [b.spectral(f) for f in (shrt.OMEGA, shrt.RHO, shrt.PHI)]
    assert not p.net.correct(b, p.scale_in, sigma, p.scale_out, dt, 0.0)
    for f, x in zip((shrt.OMEGA, shrt.RHO, shrt.PHI), before):
        assert np.array_equal(b.spectral(f).view(np.uint8), x.view(np.uint8))
    # a non-finite state gives non-finite network outputs: no correction
    bad = before[0].copy()
    bad[1, 1] = np.nan
    b.set_spectral(shrt.OMEGA, bad)
    rho_before = b.spectral(shrt.RHO)
    assert not p.net.correct(b, p.scale_in, sigma, p.scale_out, dt, 10.0)
    assert np.array_equal(b.spectral(shrt.RHO).view(np.uint8), rho_before.view(np.uint8))
    with pytest.raises(shrt.ShrtError):
        p.net.correct(b, np.zeros(3), sigma, p.scale_out, dt, 10.0)
    a.close()
    b.close()
    p.net.close()


def test_time_to_error_pipeline(gpu):
    """The measurement of the full configuration (bench.py --closure with
    config5) at reduced size: 256^2 fine runs filtered on the device to 64^2,
    plain 64^2, 128^2 and 256^2 runs and the learned 64^2 run from the same
    filtered states on two held-out seeds, each timed, and the plain method's
    time at the learned run's error, after the correction strength and band
    are calibrated on a validation seed.  Checks the measurement itself: the
    calibration keeps its best candidate, finite errors, the finest plain grid
    closest to the filtered fine run, the fallback equal to the plain run when
    the calibration chose it, and a time ratio bracketed by the plain curve."""
    import importlib.util
    here = os.path.dirname(os.path.abspath(__file__))
    spec = importlib.util.spec_from_file_location("closure_pipe", os.path.join(here, "..", "shearing-rt_amd",
                                                                               "closure.py"))
    cl = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(cl)
    print("time-to-error pipeline 256^2 -> 64^2", flush=True)
    cfg = cl.Config(n_fine=256, n_coarse=64, n_mid=128, delt=5e-4, interval=20, spinup=400, snaps=20, nu=1e-4,
                    rho0_prime=-1.0, amp=1.0, kmin=1.0, kmax_ic=16.0, kmax_net=8, width=16, layers=2, cfl=0.5,
                    kcut=8, guard=0.1)
    p = cl.Pipeline(shrt, cfg)
    x, y = p.generate(range(100, 108))
    p.fit(x, y, iters=200, batch=8, lr=5e-4)
    # correction strength and band from rollouts on a validation seed; alpha 0 is the plain solver
    table = p.calibrate([800], 400, alphas=(0.5, 1.0), kcuts=[4, 8])
    assert len(table) == 5 and table[0]["alpha"] == 0.0
    assert all(np.isfinite(t["enstrophy_spectrum"]) for t in table)
    best = min(table, key=lambda t: (t["enstrophy_spectrum"], t["alpha"]))
    assert (cfg.alpha, cfg.kcut) == (best["alpha"], best["kcut"])
    print("  calibration", [(t["alpha"], t["kcut"], round(t["enstrophy_spectrum"], 4)) for t in table], flush=True)
    res = p.evaluate([900, 901], 400, grids=[128, 256], host_path=False)
    runs = ["plain", "plain_128", "plain_256", "closure_device"]
    for r in res:
        for k in runs:
            assert np.isfinite(r[k]["enstrophy_spectrum"]) and np.isfinite(r[k]["rho"]) and r["wall_s"][k] > 0, k
        assert r["plain_256"]["enstrophy_spectrum"] < r["plain"]["enstrophy_spectrum"]
        if cfg.alpha == 0:   # the fallback: the learned run is the plain coarse run
            assert r["closure_device"] == r["plain"]
        print("  seed", r["seed"], {k: round(r[k]["enstrophy_spectrum"], 4) for k in runs},
              {k: round(r["wall_s"][k] * 1e3, 2) for k in runs})
    mean = lambda k: float(np.mean([r[k]["enstrophy_spectrum"] for r in res]))
    wall = lambda k: float(np.mean([r["wall_s"][k] for r in res]))
    pts = [(wall(k), mean(k)) for k in runs[:3]]
    t_eq, how = cl.time_at_error(pts, mean("closure_device"))
    assert how in ("interpolated", "below", "above", "nearest")
    assert min(w for w, _ in pts) <= t_eq <= max(w for w, _ in pts)
    print(f"  plain time at the learned error: {t_eq * 1e3:.2f} ms ({how}); learned {wall('closure_device') * 1e3:.2f} ms")
    p.net.close()
