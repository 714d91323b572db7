"""Learned closure pipeline: training data from the solver's own fine runs, the
spectral-convolution network trained on the device (shrt_closure_*), and the
coarse solver corrected by it, with the plain coarse and a finer plain solver
as the fallbacks it is measured against.

Data.  A fine run (n_fine^2, FP32) is filtered to the coarse grid by keeping
the coarse grid's retained modes (the same kx, ky, and the same shift since
the shift depends on kx and time only).  For snapshots F_j taken every
`interval` steps, the coarse solver restarted from F_j (Euler, AB2, AB3
start-up) gives C_j after `interval` steps; the target is the coarse
tendency defect (F_{j+1} - C_j) / (interval * delt) of omega and rho, and the
inputs are the physical omega, rho, phi of F_j plus the shear rate.

Use.  Every `interval` steps the corrected coarse run adds
alpha * interval * delt * net(state) to omega and rho, on the modes with
|kx|, ky <= kcut.  The correction is skipped (fallback to the plain coarse
step) when the network output is not finite or larger than `guard` times the
field it corrects.  The strength alpha and the band kcut are chosen after
training by rollouts on the validation trajectories (`calibrate`): the network
is fitted to one-interval defects, and how much of its prediction helps over
thousands of steps is a property of the rollout, not of the fit; alpha = 0 is
among the candidates, so the calibrated run falls back to the plain solver
when no correction helps on the validation seeds.

Error measure.  Against the filtered fine run: rho relative L2 on the coarse
modes and the relative L2 of the shell-summed enstrophy spectrum
Z(k) = 1/2 sum |omega|^2 (Hermitian-weighted) over the coarse shells.
"""
import time

import numpy as np

TWO_PI = 2 * np.pi


class Config:
    """One experiment.  Defaults: 128^2 fine, 32^2 coarse, RT-unstable
    stratification sheared at sigma ~ U(0.5, 1.5), seeded rho perturbation on
    1 <= |k| <= 4 grown into the nonlinear regime before snapshots start."""

    def __init__(self, **kw):
        self.n_fine = 128
        self.n_coarse = 32
        self.n_mid = 64              # the finer plain fallback
        self.delt = 2e-3
        self.interval = 10           # steps between snapshots / closure applications (even)
        self.spinup = 1000           # fine steps before the first snapshot
        self.snaps = 60              # snapshots per trajectory
        self.nu = 1e-3
        self.rho0_prime = 1.0        # > 0 unstable (Rayleigh-Taylor), < 0 stable stratification
        self.cfl = 5.0
        self.amp, self.kmin, self.kmax_ic, self.expo = 0.05, 1.0, 4.0, -1.0
        self.sigma_range = (0.5, 1.5)
        self.kmax_net, self.width, self.layers = 10, 32, 4
        self.guard = 0.5
        self.kcut = 0                # > 0: corrections kept to |kx|, ky <= kcut
        self.alpha = 1.0             # strength of the applied correction (0: the plain coarse solver)
        for k, v in kw.items():
            if not hasattr(self, k):
                raise KeyError(k)
            setattr(self, k, v)

    def params(self, sigma):
        return dict(Lx=TWO_PI, Ly=TWO_PI, delt=self.delt, cfl_num=self.cfl, nu=self.nu, kappa=self.nu,
                    sigma=float(sigma), g=1.0, rho0=1.0, rho0_prime=self.rho0_prime, rho_eps1=1e-2, noise=0,
                    linear=0)

    def sigma_of(self, seed):
        lo, hi = self.sigma_range
        return lo + (hi - lo) * np.random.default_rng(seed + 7919).random()


def coarse_rows(s_fine, s_coarse):
    """Rows of the fine spectrum holding the coarse grid's kx values (storage
    order: kx = i for i < nkxh, kx = -(nkx - i) above)."""
    nkx_c, nkxh_c, nkx_f = s_coarse["nkx"], s_coarse["nkxh"], s_fine["nkx"]
    return np.array([i if i < nkxh_c else nkx_f - (nkx_c - i) for i in range(nkx_c)])


class Filter:
    def __init__(self, s_fine, s_coarse):
        self.rows = coarse_rows(s_fine, s_coarse)
        self.nky = s_coarse["nky"]

    def __call__(self, F):
        return np.ascontiguousarray(F[self.rows, :self.nky])


def enstrophy_spectrum(om, kx_int, ky_int):
    """Z(k) = 1/2 sum over the shell round(|k|) = k of |omega|^2, counting the
    conjugate half plane (ky > 0 twice)."""
    w = np.where(ky_int[None, :] > 0, 2.0, 1.0)
    kk = np.rint(np.sqrt(kx_int[:, None] ** 2 + ky_int[None, :] ** 2)).astype(int)
    return np.bincount(kk.ravel(), weights=(0.5 * w * np.abs(om) ** 2).ravel())


def k_ints(s):
    nkx, nky, nkxh = s["nkx"], s["nky"], s["nkxh"]
    kx = np.array([i if i < nkxh else i - nkx for i in range(nkx)])
    return kx, np.arange(nky)


class Pipeline:
    def __init__(self, shrt, cfg):
        self.shrt, self.cfg = shrt, cfg
        self.sf = shrt.sizes(cfg.n_fine, cfg.n_fine)
        self.sc = shrt.sizes(cfg.n_coarse, cfg.n_coarse)
        self.filt = Filter(self.sf, self.sc)
        self.net = None
        self.scale_in = np.ones(3)
        self.scale_out = np.ones(2)

    def ctx(self, n, sigma, graph_steps=None):
        gs = self.cfg.interval if graph_steps is None else graph_steps
        return self.shrt.Shrt(nx=n, ny=n, precision=self.shrt.FP32, graph_steps=gs, **self.cfg.params(sigma))

    def fine_snapshots(self, seed, count):
        """Filtered (omega, rho, time, shift) of a fine run every `interval` steps."""
        cfg, S = self.cfg, self.shrt
        f = self.ctx(cfg.n_fine, cfg.sigma_of(seed))
        f.initialize_seeded(seed=seed, kmin=cfg.kmin, kmax=cfg.kmax_ic, amp=cfg.amp, expo=cfg.expo)
        if cfg.spinup:
            f.time_advance(cfg.spinup)
        out = []
        nc = cfg.n_coarse
        for j in range(count):
            if j:
                f.time_advance(cfg.interval)
            st = f.state()
            sh, _ = f.shift()
            # the filter runs on the device: only the coarse modes cross to the host
            out.append((f.spectral_modes(S.OMEGA, nc).astype(np.complex128),
                        f.spectral_modes(S.RHO, nc).astype(np.complex128), st["time"],
                        np.ascontiguousarray(sh[self.filt.rows]), st["halvings"]))
        f.close()
        return out

    def load_state(self, c, om, rh, t, sh):
        c.set_time(0, t, self.cfg.delt)
        c.set_shift(sh)
        c.set_spectral(self.shrt.OMEGA, om)
        c.set_spectral(self.shrt.RHO, rh)

    def inputs_of(self, c, sigma):
        S = self.shrt
        om, rh, ph = c.spectral(S.OMEGA), c.spectral(S.RHO), c.spectral(S.PHI)
        x = np.empty((4, self.cfg.n_coarse ** 2))
        for i, F in enumerate((om, rh, ph)):
            x[i] = c.ktox(F).ravel() / self.scale_in[i]
        x[3] = sigma
        return x

    def generate(self, seeds, log=None):
        """Inputs [S][4][n^2] (unscaled) and targets [S][2][n^2], float32."""
        cfg, S = self.cfg, self.shrt
        seeds = list(seeds)
        npts = cfg.n_coarse ** 2
        x = np.empty((len(seeds) * cfg.snaps, 4, npts), dtype=np.float32)
        y = np.empty((len(seeds) * cfg.snaps, 2, npts), dtype=np.float32)
        k = 0
        saved = self.scale_in
        self.scale_in = np.ones(3)
        for si, seed in enumerate(seeds):
            sigma = cfg.sigma_of(seed)
            snaps = self.fine_snapshots(seed, cfg.snaps + 1)
            self.max_halvings = max(getattr(self, "max_halvings", 0), max(sn[4] for sn in snaps))
            c = self.ctx(cfg.n_coarse, sigma)
            for j in range(cfg.snaps):
                om, rh, t, sh, _ = snaps[j]
                self.load_state(c, om, rh, t, sh)
                xj = self.inputs_of(c, sigma)
                c.time_advance(cfg.interval)
                dt_tot = cfg.interval * cfg.delt
                d_om = (snaps[j + 1][0] - c.spectral(S.OMEGA)) / dt_tot
                d_rh = (snaps[j + 1][1] - c.spectral(S.RHO)) / dt_tot
                x[k] = xj
                y[k, 0] = c.ktox(d_om).ravel()
                y[k, 1] = c.ktox(d_rh).ravel()
                k += 1
            c.close()
            if log is not None and (si % 8 == 7 or si == len(seeds) - 1):
                log(f"closure data: {si + 1}/{len(seeds)} trajectories")
        self.scale_in = saved
        x, y = x[:k], y[:k]
        ok = np.isfinite(x).all(axis=(1, 2)) & np.isfinite(y).all(axis=(1, 2))
        self.dropped = int((~ok).sum())   # samples from runs that blew up
        return (x, y) if self.dropped == 0 else (x[ok], y[ok])

    def normalise(self, x, y=None):
        xn = x.copy()
        xn[:, :3] /= self.scale_in[None, :, None].astype(x.dtype)
        return xn if y is None else (xn, y / self.scale_out[None, :, None].astype(y.dtype))

    def predict(self, xn):
        """Network outputs (normalised units) for normalised inputs, in batches."""
        B = self.net.max_batch
        return np.concatenate([self.net.forward(xn[i:i + B].astype(np.float32)) for i in range(0, len(xn), B)])

    def r2(self, xn, yn):
        """Fraction of the target variance explained, per output channel."""
        e = ((self.predict(xn) - yn) ** 2).mean(axis=(0, 2))
        return 1.0 - e / (yn ** 2).mean(axis=(0, 2))

    def fit(self, x, y, iters=2000, batch=8, lr=1e-3, seed=1, val=None, chunk=250):
        """Train the closure network on (x, y); returns the loss history.  With a
        validation set val = (xv, yv) training runs in chunks and keeps the
        parameters with the best mean validation R^2 (early stopping)."""
        cfg = self.cfg
        self.nsamples = len(x)
        self.scale_in = np.maximum(x[:, :3].std(axis=(0, 2)), 1e-30)
        self.scale_out = np.maximum(y.std(axis=(0, 2)), 1e-30)
        xn, yn = self.normalise(x, y)
        if self.net is None:
            self.net = self.shrt.Closure(cfg.n_coarse, cfg.kmax_net, cfg.width, cfg.layers, max_batch=batch,
                                         seed=seed)
        self.net.set_data(xn.astype(np.float32, copy=False), yn.astype(np.float32, copy=False))
        if val is None:
            return self.net.train(iters, batch, lr, seed=seed)
        xv, yv = self.normalise(*val)
        best, best_p, losses, self.val_curve = -np.inf, None, [], []
        done = 0
        while done < iters:
            k = min(chunk, iters - done)
            losses.append(self.net.train(k, batch, lr, seed=seed + done))
            done += k
            score = float(np.mean(self.r2(xv, yv)))
            self.val_curve.append((done, score))
            if np.isfinite(score) and score > best:
                best, best_p = score, self.net.params()
        if best_p is not None:
            self.net.set_params(best_p)
        self.val_best = best
        return np.concatenate(losses)

    def correction(self, c, sigma):
        """Spectral corrections of omega and rho (per unit time), or None for the fallback."""
        out = self.net.forward(self.inputs_of(c, sigma)[None].astype(np.float32))[0].astype(np.float64)
        if not np.all(np.isfinite(out)):
            return None
        n = self.cfg.n_coarse
        corr = [c.xtok((out[i] * self.scale_out[i]).reshape(n, n)) for i in range(2)]
        if self.cfg.kcut > 0:
            kx, ky = k_ints(self.sc)
            keep = (np.abs(kx)[:, None] <= self.cfg.kcut) & (ky[None, :] <= self.cfg.kcut)
            for F in corr:
                F[~keep] = 0
        return corr

    def run_coarse(self, om, rh, t, sh, sigma, steps, closure=False, n=None, device=False):
        """A coarse (or, with n, another plain) run from a filtered state; returns
        the final (omega, rho) restricted to the coarse modes, the wall time and
        the number of corrections skipped by the fallback.  device=True applies
        the corrections with shrt_closure_correct (no host round trip), the same
        arithmetic as the host path below."""
        cfg, S = self.cfg, self.shrt
        n = n or cfg.n_coarse
        c = self.ctx(n, sigma)
        if n == cfg.n_coarse:
            self.load_state(c, om, rh, t, sh)
        else:
            s_n = self.shrt.sizes(n, n)
            rows = coarse_rows(s_n, self.sc)
            Om = np.zeros(c.shape_k, dtype=np.complex64)
            Rh = np.zeros_like(Om)
            Om[rows, :self.sc["nky"]] = om
            Rh[rows, :self.sc["nky"]] = rh
            # the shift of a row is kx times that of kx = 1 (Lx = 2 pi); the coarse rows keep theirs
            kxi, _ = k_ints(s_n)
            shn = kxi * sh[1]
            shn[rows] = sh
            self.load_state(c, Om, Rh, t, shn)
        skipped = 0
        closure = closure and cfg.alpha > 0   # strength 0: the plain solver
        c.synchronize()
        t0 = time.perf_counter()
        done = 0
        if not closure:                # plain: one call, captured batches back to back
            c.time_advance(steps)
            done = steps
        while done < steps:
            k = min(cfg.interval, steps - done)
            if closure and device and k == cfg.interval:
                if not self.net.correct(c, self.scale_in, sigma, self.scale_out, cfg.alpha * cfg.interval * cfg.delt,
                                        cfg.guard, cfg.kcut):
                    skipped += 1
            elif closure and k == cfg.interval:
                corr = self.correction(c, sigma)
                om_c, rh_c = c.spectral(S.OMEGA), c.spectral(S.RHO)
                dt_tot = cfg.alpha * cfg.interval * cfg.delt
                if corr is None or (np.linalg.norm(corr[0]) * dt_tot > cfg.guard * np.linalg.norm(om_c) or
                                    np.linalg.norm(corr[1]) * dt_tot > cfg.guard * np.linalg.norm(rh_c)):
                    skipped += 1
                else:
                    c.set_spectral(S.OMEGA, om_c + dt_tot * corr[0])
                    c.set_spectral(S.RHO, rh_c + dt_tot * corr[1])
            c.time_advance(k)
            done += k
        wall = time.perf_counter() - t0
        if n == cfg.n_coarse:
            out = (c.spectral(S.OMEGA), c.spectral(S.RHO))
        else:
            f = Filter(self.shrt.sizes(n, n), self.sc)
            out = (f(c.spectral(S.OMEGA)), f(c.spectral(S.RHO)))
        c.close()
        return out, wall, skipped

    def calibrate(self, seeds, steps, alphas=(0.0, 0.25, 0.5, 1.0), kcuts=None, log=None):
        """Choose the correction strength alpha and band kcut by rollouts on
        (validation) seeds: every candidate runs `steps` corrected coarse steps
        from each seed's first filtered snapshot, and the pair with the lowest
        mean enstrophy-spectrum error against the filtered fine run is kept in
        the configuration (alpha = 0 is the plain coarse solver).  Returns the
        table of mean errors per candidate."""
        cfg = self.cfg
        kcuts = [cfg.kcut] if kcuts is None else list(kcuts)
        cands = [(0.0, cfg.kcut)] + [(a, k) for a in alphas if a > 0 for k in kcuts]
        starts = []
        for seed in seeds:
            snaps = self.fine_snapshots(seed, steps // cfg.interval + 1)
            starts.append((seed, cfg.sigma_of(seed), snaps[0], snaps[-1][:2]))
        saved = (cfg.alpha, cfg.kcut)
        table = []
        for a, k in cands:
            cfg.alpha, cfg.kcut = a, k
            errs = []
            for seed, sigma, (om, rh, t, sh, _), ref in starts:
                out, _, _ = self.run_coarse(om, rh, t, sh, sigma, steps, closure=True, device=True)
                errs.append(self.errors(out, ref))
            row = dict(alpha=a, kcut=k, enstrophy_spectrum=float(np.mean([e["enstrophy_spectrum"] for e in errs])),
                       rho=float(np.mean([e["rho"] for e in errs])))
            table.append(row)
            if log is not None:
                log(f"closure calibration: alpha {a} kcut {k}: enstrophy spectrum {row['enstrophy_spectrum']:.4f}, "
                    f"rho {row['rho']:.4f}")
        cfg.alpha, cfg.kcut = saved
        best = min(table, key=lambda r: (r["enstrophy_spectrum"], r["alpha"]))
        cfg.alpha, cfg.kcut = best["alpha"], best["kcut"]
        return table

    def errors(self, run, ref):
        kx, ky = k_ints(self.sc)
        zr = enstrophy_spectrum(ref[0], kx, ky)
        zg = enstrophy_spectrum(run[0], kx, ky)
        m = min(len(zr), self.cfg.kmax_net + 1)
        ens = float(np.linalg.norm(zg[1:m] - zr[1:m]) / np.linalg.norm(zr[1:m]))
        rho = float(np.linalg.norm(run[1] - ref[1]) / np.linalg.norm(ref[1]))
        return dict(enstrophy_spectrum=ens, rho=rho)

    def evaluate(self, seeds, steps, grids=None, host_path=True, log=None):
        """Held-out seeds: plain coarse, corrected coarse (device path, and the
        host path when host_path) and plain runs on the finer grids against the
        filtered fine run after `steps` steps from the first snapshot; per run
        the error measures and the wall time of the run (GPU work and launches,
        all from the same filtered state, same step size)."""
        grids = [self.cfg.n_mid] if grids is None else list(grids)
        res = []
        for seed in seeds:
            sigma = self.cfg.sigma_of(seed)
            snaps = self.fine_snapshots(seed, steps // self.cfg.interval + 1)
            om, rh, t, sh, _ = snaps[0]
            ref = snaps[-1][:2]
            plain, w_plain, _ = self.run_coarse(om, rh, t, sh, sigma, steps)
            dev, w_dev, skipped_dev = self.run_coarse(om, rh, t, sh, sigma, steps, closure=True, device=True)
            r = dict(seed=seed, sigma=sigma, plain=self.errors(plain, ref), closure_device=self.errors(dev, ref),
                     skipped_device=skipped_dev, wall_s=dict(plain=w_plain, closure_device=w_dev))
            if host_path:
                corr, w_corr, skipped = self.run_coarse(om, rh, t, sh, sigma, steps, closure=True)
                r.update(closure=self.errors(corr, ref), skipped=skipped,
                         device_vs_host=float(np.linalg.norm(dev[0] - corr[0]) / np.linalg.norm(corr[0])))
                r["wall_s"]["closure"] = w_corr
            for n in grids:
                out, w, _ = self.run_coarse(om, rh, t, sh, sigma, steps, n=n)
                key = "mid" if (n == self.cfg.n_mid and len(grids) == 1) else f"plain_{n}"
                r[key] = self.errors(out, ref)
                r["wall_s"][key] = w
            res.append(r)
            if log is not None:
                log(f"closure eval seed {seed}: " + ", ".join(
                    f"{k} {r[k]['enstrophy_spectrum']:.3f}/{r[k]['rho']:.3f} ({r['wall_s'][k] * 1e3:.1f} ms)"
                    for k in r["wall_s"]))
        return res


def time_at_error(points, err):
    """GPU time the plain method needs for error `err`, from its (time, error)
    points by log-log interpolation between the two grids that bracket it.
    Returns (time, bracket): bracket is "interpolated", "below" (err under every
    plain error: the finest plain time is a lower bound) or "above" (err over
    every plain error: the coarsest plain time is an upper bound)."""
    pts = sorted(points)                  # by time
    ts = np.array([p[0] for p in pts])
    es = np.array([p[1] for p in pts])
    if err <= es.min():
        return float(ts[np.argmin(es)]), "below"
    if err >= es[0]:
        return float(ts[0]), "above"
    for i in range(len(pts) - 1):
        e0, e1 = es[i], es[i + 1]
        if (e0 - err) * (e1 - err) <= 0 and e0 != e1:
            w = (np.log(err) - np.log(e0)) / (np.log(e1) - np.log(e0))
            return float(np.exp(np.log(ts[i]) + w * (np.log(ts[i + 1]) - np.log(ts[i])))), "interpolated"
    return float(ts[np.argmin(np.abs(es - err))]), "nearest"
