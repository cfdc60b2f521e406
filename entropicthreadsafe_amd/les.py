"""Closure LES: the learned subgrid closure judged against the plain method.

A fine two-component run (emulsion in linearly forced turbulence) supplies
box-filtered snapshots for training the closure (tslb_closure_add_samples,
tslb_closure_train, all on the GPU).  From the filtered fine state at time
t0, two coarse runs follow the flow for the same physical time (acoustic
scaling: dx and dt both grow by the filter factor r): the plain coarse
lattice and the same lattice with the trained closure attached.  Each is
compared with the filtered fine run at the same time through the shell-
averaged kinetic-energy spectrum E(k) over the resolved band, and the wall
time of the closure run is compared with the fine run's.

Parameter scaling to the coarse grid (lattice units): nu_c = nu_f / r,
sigma_c = sigma_f / r, linear-forcing rate A_c = r A_f; phi and velocities
carry over.  The initial velocity is a seeded solenoidal random field with
spectrum E(k) ~ k^4 exp(-2 (k/k0)^2) (host FFTs), plus the droplets of the
emulsion workload.
"""
import json
import time

import numpy as np

from . import configs
from .tslb import Closure, Lattice


def solenoidal_field(n, k0, urms, seed, workers=8):
    """Random divergence-free velocity (3, n, n, n) float32 with spectrum
    ~ k^4 exp(-2 (k/k0)^2), scaled to the given rms per component (host FFTs)."""
    import scipy.fft
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(n, 1.0 / n).astype(np.float32)
    kr = np.fft.rfftfreq(n, 1.0 / n).astype(np.float32)
    kz, ky, kx = k1[:, None, None], k1[None, :, None], kr[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    kk = np.sqrt(k2)
    amp = (kk * kk * np.exp(-(kk / k0) ** 2) / np.maximum(kk, 1.0)).astype(np.float32)  # sqrt(E/(4 pi k^2))
    amp[0, 0, 0] = 0.0
    comps = []
    for _ in range(3):
        c = rng.standard_normal(amp.shape, dtype=np.float32) + 1j * rng.standard_normal(amp.shape, dtype=np.float32)
        comps.append((c * amp).astype(np.complex64))
    # project out the compressive part: a - k (k.a)/k^2
    dot = (kx * comps[0] + ky * comps[1] + kz * comps[2]) / np.maximum(k2, 1.0)
    out = np.empty((3, n, n, n), np.float32)
    for a, (c, kc) in enumerate(zip(comps, (kx, ky, kz))):
        out[a] = scipy.fft.irfftn(c - kc * dot, s=(n, n, n), workers=workers)
    del comps, dot
    out *= urms / np.sqrt(np.mean(out.astype(np.float64) ** 2))
    return out


def box(a, r):
    """r^3 block means of a (nz, ny, nx)."""
    nz, ny, nx = a.shape
    return a.reshape(nz // r, r, ny // r, r, nx // r, r).mean(axis=(1, 3, 5), dtype=np.float64)


def filtered_state(lat, r):
    """rho_bar, Favre velocity and phi_bar of a lattice (host, float64)."""
    rho = lat.download_field("rho").astype(np.float64)
    rb = box(rho, r)
    vel = [box(rho * lat.download_field(c), r) / rb for c in ("u", "v", "w")]
    phi = box(lat.download_field("phi"), r)
    return rb, vel, phi


def energy_spectrum(u, v, w):
    """Shell-averaged kinetic-energy spectrum E(k), k = 0 .. n/2."""
    n = u.shape[0]
    e = np.zeros(n // 2 + 1)
    k1 = np.fft.fftfreq(n, 1.0 / n)
    kr = np.fft.rfftfreq(n, 1.0 / n)
    kz, ky, kx = np.meshgrid(k1, k1, kr, indexing="ij")
    kk = np.rint(np.sqrt(kx * kx + ky * ky + kz * kz)).astype(int)
    wgt = np.where(kx == 0, 1.0, 2.0)   # rfft half-space
    if n % 2 == 0:
        wgt = np.where(kx == n // 2, 1.0, wgt)
    tot = np.zeros_like(kk, dtype=np.float64)
    for a in (u, v, w):
        f = np.fft.rfftn(a) / a.size
        tot += 0.5 * wgt * (f.real ** 2 + f.imag ** 2)
    m = kk <= n // 2
    np.add.at(e, kk[m], tot[m])
    return e


def spectrum_error(e, ref, kmax):
    """Mean relative deviation of E(k) from the reference over 1 <= k <= kmax."""
    k = np.arange(1, kmax + 1)
    return float(np.mean(np.abs(e[k] / ref[k] - 1.0)))


def _coarse_lattice(nc, kw, r, rb, vel, phi):
    ckw = dict(kw)
    ckw["tau"] = 0.5 + (kw["tau"] - 0.5) / r
    ckw["sigma"] = kw["sigma"] / r
    ckw["lin_force"] = kw["lin_force"] * r
    lat = Lattice(nc, nc, nc, **ckw)
    lat.upload_field("rho", rb.astype(np.float32))
    for c, a in zip(("u", "v", "w"), vel):
        lat.upload_field(c, a.astype(np.float32))
    lat.upload_field("phi", phi.astype(np.float32))
    lat.init_equilibrium(True)
    return lat


def _timed_steps(lat, steps):
    lat.sync()
    t = time.perf_counter()
    lat.step(steps)
    lat.sync()
    return time.perf_counter() - t


def run(n=512, r=4, k0=4.0, urms=0.04, spinup=1500, snaps=4, snap_every=250, horizon=2000,
        train_steps=9000, batch=16384, lr=2e-3, seed=1, drops=True, balance=False, log=print):
    """The experiment; returns a dict of its results.  drops=False runs the
    same two-component lattice without droplets (phi = 0): single-phase
    turbulence, a check of the coarse-grid scaling.  The mean subgrid
    dissipation of the data and of the trained model is always recorded;
    balance=True adds a coarse run with nu_t scaled to equalise them."""
    case = configs.emulsion(n=n)
    kw = dict(case.lattice)
    nc = n // r
    phi0 = case.fields["phi"] if drops else np.zeros_like(case.fields["phi"])
    res = dict(workload=f"closure-les {'emulsion' if drops else 'single-phase'}-2c {n}^3 -> {nc}^3",
               n_fine=n, n_coarse=nc, factor=r, spinup=spinup, horizon_fine_steps=horizon,
               horizon_coarse_steps=horizon // r, notes=case.notes if drops else "no droplets")
    vel0 = solenoidal_field(n, k0, urms, seed)
    with Lattice(n, n, n, **kw) as fine, Closure(max_samples=snaps * nc ** 3) as model:
        fine.upload_field("rho", case.fields["rho"])
        for c, a in zip(("u", "v", "w"), vel0):
            fine.upload_field(c, a)
        fine.upload_field("phi", phi0)
        fine.init_equilibrium(True)
        del vel0
        t_spin = _timed_steps(fine, spinup)
        log(f"les: fine spin-up {spinup} steps in {t_spin:.1f} s")
        for s in range(snaps):
            fine.step(snap_every)
            model.add_samples(fine, r)
        log(f"les: {model.num_samples} training samples from {snaps} snapshots")
        rb, vel, phi = filtered_state(fine, r)     # coarse initial state, time t0
        t_fine = _timed_steps(fine, horizon)
        frb, fvel, _ = filtered_state(fine, r)
        e_ref = energy_spectrum(*fvel)
        e_t0 = energy_spectrum(*vel)
        res["fine_seconds"] = t_fine
        log(f"les: fine horizon {horizon} steps in {t_fine:.2f} s")

        # training: inputs scaled to unit spread, outputs to the target size
        feat, aux = model.get_samples()
        fstd = feat.std(axis=1) + 1e-12
        t_rms = float(np.sqrt((aux[1:] ** 2).mean()))
        t_scale = t_rms or 1.0   # r = 1: no subgrid stress to fit
        theta0 = _init_theta(seed, 1.0 / fstd, t_rms or 1e-9)
        model.set_params(theta0)
        l0 = model.loss_grad(t_scale, grad=False)
        t0 = time.perf_counter()
        # three stages of decreasing learning rate
        for stage, f in enumerate((1.0, 0.3, 0.1)):
            model.train(train_steps // 3, batch, lr * f, t_scale, seed=seed + stage)
        t_train = time.perf_counter() - t0
        l1 = model.loss_grad(t_scale, grad=False)
        res.update(train_loss_initial=l0, train_loss_final=l1, train_seconds=t_train, train_steps=train_steps,
                   batch=batch, t_scale=t_scale, samples=model.num_samples)
        log(f"les: training loss {l0:.4g} -> {l1:.4g} ({train_steps} Adam steps in {t_train:.1f} s)")

        # energy balance over the samples: the factor on nu_t that makes the
        # model's mean subgrid dissipation <2 rho nu_t S:S - dsigma:S> equal
        # the filtered data's <-tau:S> (off-diagonal pairs counted twice)
        out = model.forward().astype(np.float64)
        S = feat[:6].astype(np.float64)
        c = np.array([1.0, 1.0, 1.0, 2.0, 2.0, 2.0])[:, None]
        eps_true = float(-(c * aux[1:].astype(np.float64) * S).sum(0).mean())
        eps_nu = float((2.0 * aux[0].astype(np.float64) * out[0] * (c * S * S).sum(0)).mean())
        eps_ds = float(-(c * out[1:] * S).sum(0).mean())
        alpha = (eps_true - eps_ds) / eps_nu if eps_nu > 0 else 0.0
        del out, S
        res.update(eps_subgrid=eps_true, eps_model=eps_nu + eps_ds, nu_t_balance=alpha)
        log(f"les: subgrid dissipation {eps_true:.4g}, model {eps_nu + eps_ds:.4g}; nu_t factor {alpha:.4g}")

        # the coarse runs
        spectra = {}
        theta = model.get_params()
        names = ("plain", "closure") + (("balanced",) if balance and alpha > 0 else ())
        for name in names:
            if name == "balanced":
                tb = theta.copy()
                tb[-7] *= alpha            # the nu_t output scale
                model.set_params(tb)
            with _coarse_lattice(nc, kw, r, rb, vel, phi) as lat:
                if name != "plain":
                    lat.set_closure(model)
                    clo = lat.download_closure()   # fields of the t0 state
                    res[f"{name}_nu_t_mean"] = float(clo[0].mean())
                    res[f"{name}_nu_t_max"] = float(clo[0].max())
                    res[f"{name}_dsigma_rms"] = float(np.sqrt((clo[1:].astype(np.float64) ** 2).mean()))
                    res["nu_coarse"] = (0.5 + (kw["tau"] - 0.5) / r - 0.5) / 3.0
                dt = _timed_steps(lat, horizon // r)
                m = lat.download()
                if name != "plain":
                    lat.set_closure(None)
            ok = bool(np.isfinite(m).all())
            spectra[name] = energy_spectrum(m[1], m[2], m[3]) if ok else None
            res[f"{name}_seconds"] = dt
            res[f"{name}_finite"] = ok
            log(f"les: coarse {name} {horizon // r} steps in {dt:.3f} s, finite={ok}")
    kmax = nc // 4
    res["kmax"] = kmax
    # how far the flow moved over the horizon: the t0 spectrum against the end
    res["persistence_spectrum_error"] = spectrum_error(e_t0, e_ref, kmax)
    for name, e in spectra.items():
        res[f"{name}_spectrum_error"] = spectrum_error(e, e_ref, kmax) if e is not None else None
    res["speedup_closure_vs_fine"] = res["fine_seconds"] / res["closure_seconds"]
    res["spectrum_k"] = list(range(kmax + 1))
    res["spectrum_fine_filtered"] = [float(x) for x in e_ref[:kmax + 1]]
    for name, e in spectra.items():
        res[f"spectrum_{name}"] = [float(x) for x in e[:kmax + 1]] if e is not None else None
    return res


def _init_theta(seed, in_scale, t_scale):
    """Random MLP weights (He), zero biases; nu_t output scale 1e-4 and the
    stress corrections scaled to the target size."""
    rng = np.random.default_rng(seed)
    parts = []
    for i, o in ((10, 64), (64, 64), (64, 64), (64, 7)):
        parts += [(rng.standard_normal((i, o)) * np.sqrt(2.0 / i)).ravel(), np.zeros(o)]
    parts[-2] = parts[-2] * 0.1            # small initial output layer
    parts += [np.asarray(in_scale, np.float64), np.array([1e-4] + [t_scale] * 6)]
    return np.concatenate(parts).astype(np.float32)


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--factor", type=int, default=4)
    ap.add_argument("--spinup", type=int, default=1500)
    ap.add_argument("--horizon", type=int, default=2000)
    ap.add_argument("--train-steps", type=int, default=9000)
    ap.add_argument("--no-drops", action="store_true")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    out = run(n=a.n, r=a.factor, spinup=a.spinup, horizon=a.horizon, train_steps=a.train_steps,
              drops=not a.no_drops)
    js = json.dumps(out)
    if a.out:
        with open(a.out, "w") as f:
            f.write(js + "\n")
    print(json.dumps({k: v for k, v in out.items() if not k.startswith("spectrum")}))
