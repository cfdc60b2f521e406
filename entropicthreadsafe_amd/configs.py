"""Workload set-up: synthetic initial states of the TSLB configurations.

Each builder returns a `Case`: the Lattice keyword arguments, the initial
rho/u/v/w/phi fields (numpy, shape (nz, ny, nx)), an optional isfluid map and a
list of seeded-noise terms `(field, amplitude, l)` meaning
field += amp * (2 r - 1), r = rand_noseeded(i, j, k, l) at 1-based (i, j, k).
Noise is applied on the device by `apply()` (tslb_seed_noise) or on the host by
`host_fields()`; both give the same values.

Configurations (all seeds through the reference's counter-based generator,
TSLB_newBCS/vars_single_3D_module.f90:102-165):
  tgv       3D Taylor-Green vortex, one component, periodic
  droplet   static droplet, two components, periodic
  channel   pressure-driven channel, one component, walls at y, additive force
  emulsion  droplets in linearly forced turbulence, two components, periodic
"""
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np


# ---------------------------------------------------------------------------
# vectorised counter-based generator (host side, small and medium grids)
# ---------------------------------------------------------------------------
def rand_noseeded_grid(nx, ny, nz, l):
    """rand_noseeded(i, j, k, l) for every 1-based node, shape (nz, ny, nx)."""
    k, j, i = np.meshgrid(np.arange(1, nz + 1), np.arange(1, ny + 1), np.arange(1, nx + 1),
                          indexing="ij")
    isub = (i % 178 + 1).astype(np.int64)
    jsub = (j % 178 + 1).astype(np.int64)
    ksub = (k % 178 + 1).astype(np.int64)
    lsub = int(l) % 169
    out = []
    for _ in range(2):
        s = np.zeros(isub.shape, dtype=np.float32)
        t = np.float32(0.5)
        for _ in range(24):
            msub = ((isub * jsub) % 179 * ksub) % 179
            isub, jsub, ksub = jsub, ksub, msub
            lsub = (53 * lsub + 1) % 169
            s = np.where((lsub * msub) % 64 >= 32, s + t, s).astype(np.float32)
            t = np.float32(t * np.float32(0.5))
        out.append(s)
    uni = (out[0] - out[1]).astype(np.float32)
    uni = np.where(uni < 0, uni + np.float32(1.0), uni).astype(np.float32)
    cc = np.float32(362436.0 / 16777216.0)
    cd = np.float32(7654321.0 / 16777216.0)
    cm = np.float32(16777213.0 / 16777216.0)
    csub = np.float32(cc - cd)
    if csub < 0:
        csub = np.float32(csub + cm)
    uni = (uni - csub).astype(np.float32)
    uni = np.where(uni < 0, uni + np.float32(1.0), uni).astype(np.float32)
    return uni


@dataclass
class Case:
    name: str
    shape: Tuple[int, int, int]            # (nz, ny, nx)
    lattice: Dict                           # Lattice(nx, ny, nz, **lattice) keyword arguments
    fields: Dict[str, np.ndarray]           # rho, u, v, w, phi
    noise: List[Tuple[str, float, int]] = field(default_factory=list)
    isfluid: Optional[np.ndarray] = None
    prepush_phi: bool = True
    notes: str = ""

    def host_fields(self):
        """Initial fields with the seeded noise applied on the host."""
        f = {k: v.astype(np.float32).copy() for k, v in self.fields.items()}
        nz, ny, nx = self.shape
        for name, amp, l in self.noise:
            r = rand_noseeded_grid(nx, ny, nz, l)
            f[name] = (f[name] + np.float32(amp) * (np.float32(2.0) * r - np.float32(1.0))).astype(np.float32)
        return f

    def apply(self, lat):
        """Upload the initial state to a Lattice (noise seeded on the device)."""
        if self.isfluid is not None:
            lat.set_isfluid(self.isfluid)
        for name in ("rho", "u", "v", "w", "phi"):
            lat.upload_field(name, self.fields[name])
        for name, amp, l in self.noise:
            lat.seed_noise(name, amp, l)
        lat.init_equilibrium(prepush_phi=self.prepush_phi)


def _grid(nx, ny, nz):
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return i, j, k


def tgv(n=64, u0=0.05, re=1600.0, length=None):
    """3D Taylor-Green vortex, one component, periodic n^3.

    nu = u0 L / Re with L = n (the length scale the workload names), so
    tau = 0.5 + 3 nu; x = i - 1 (0-based index), k = 2 pi / n.
    """
    L = float(n if length is None else length)
    nu = u0 * L / re
    tau = 0.5 + 3.0 * nu
    i, j, k = _grid(n, n, n)
    kk = 2.0 * np.pi / n
    u = u0 * np.sin(kk * i) * np.cos(kk * j) * np.cos(kk * k)
    v = -u0 * np.cos(kk * i) * np.sin(kk * j) * np.cos(kk * k)
    z = np.zeros_like(u)
    return Case("tgv", (n, n, n), dict(ncomp=1, tau=tau, periodic=(1, 1, 1)),
                dict(rho=np.ones_like(u), u=u, v=v, w=z, phi=z.copy()),
                notes=f"nu={nu:.6g} tau={tau:.6g}")


def droplet(n=256, radius=40.0, sigma=0.02, beta=0.7, noise=1e-3, tau=1.0):
    """Static droplet of component A (phi=1) in B, two components, periodic n^3.

    Sphere of the given radius centred at node (n/2, n/2, n/2) (1-based); seeded
    velocity noise with l = 11, 22, 33; sharpening sharp_c = 3 beta (the
    reference's 0.3*3 reads as beta/cs^2, recursive_2c_TSLB3D.f90:28).
    """
    i, j, k = _grid(n, n, n)
    c = n / 2.0
    r2 = (i + 1 - c) ** 2 + (j + 1 - c) ** 2 + (k + 1 - c) ** 2
    phi = (r2 <= radius ** 2).astype(np.float64)
    z = np.zeros_like(phi)
    return Case("droplet", (n, n, n),
                dict(ncomp=2, tau=tau, sigma=sigma, sharp_c=3.0 * beta, periodic=(1, 1, 1)),
                dict(rho=np.ones_like(phi), u=z, v=z.copy(), w=z.copy(), phi=phi),
                noise=[("u", noise, 11), ("v", noise, 22), ("w", noise, 33)])


def colour_droplet(n=32, radius=8.0, sigma=0.02, beta=0.7, tau=1.0, tau_b=1.0, noise=1e-3, width=1.0):
    """Static droplet in the colour-gradient model (ncomp 3): colour field
    psi = tanh((R - r)/width) (+1 inside, component A), seeded velocity noise
    with l = 11, 22, 33; viscosities from tau (A) and tau_b (B)."""
    i, j, k = _grid(n, n, n)
    c = n / 2.0
    r = np.sqrt((i + 1 - c) ** 2 + (j + 1 - c) ** 2 + (k + 1 - c) ** 2)
    psi = np.tanh((radius - r) / width)
    z = np.zeros_like(psi)
    return Case("colour_droplet", (n, n, n),
                dict(ncomp=3, tau=tau, tau_b=tau_b, beta=beta, sigma=sigma, periodic=(1, 1, 1)),
                dict(rho=np.ones_like(psi), u=z, v=z.copy(), w=z.copy(), phi=psi),
                noise=[("u", noise, 11), ("v", noise, 22), ("w", noise, 33)], prepush_phi=False,
                notes=f"colour droplet R={radius}")


def channel(nx=512, ny=256, nz=256, re_tau=180.0, u_tau=0.005, pert=0.05, ub_over_utau=15.6):
    """Pressure-driven channel, one component: x, z periodic, halfway
    bounce-back walls just outside y = 1 and y = ny (half height h = ny/2).

    nu = u_tau h / Re_tau, body force F_x = u_tau^2 / h (additive,
    recursive_TSLB3D.f90:571-707).  Start: parabolic profile with bulk velocity
    ub_over_utau * u_tau plus seeded perturbations of pert * U_b on u, v, w.
    """
    h = ny / 2.0
    nu = u_tau * h / re_tau
    tau = 0.5 + 3.0 * nu
    fx = u_tau ** 2 / h
    i, j, k = _grid(nx, ny, nz)
    yy = j + 0.5                      # distance from the lower wall
    ub = ub_over_utau * u_tau
    u = 1.5 * ub * (yy / h) * (2.0 - yy / h)
    z = np.zeros_like(u)
    amp = pert * ub
    return Case("channel", (nz, ny, nx),
                dict(ncomp=1, tau=tau, force=(fx, 0.0, 0.0), force_mode=0, periodic=(1, 0, 1)),
                dict(rho=np.ones_like(u), u=u, v=z, w=z.copy(), phi=z.copy()),
                noise=[("u", amp, 101), ("v", amp, 102), ("w", amp, 103)],
                notes=f"nu={nu:.6g} tau={tau:.6g} Fx={fx:.6g}")


def place_droplets(n, ndrop, rmin, rmax, seed_l=7):
    """Non-overlapping droplets by seeded rejection sampling (periodic box).

    Candidate m draws centre and radius from rand_noseeded(a, b, q, seed_l)
    with a = m mod 178 + 1, b = m div 178 + 1, q = 1..4 (the generator reduces
    each seed modulo 178, so m itself cannot be the only varying seed).
    """
    from . import tslb as _t  # the generator of the library (host C)
    drops = []
    acc = np.zeros((0, 4))
    m = 0
    while len(drops) < ndrop and m < 178 * 178:
        a, b = m % 178 + 1, m // 178 + 1
        m += 1
        cx = _t.rand_noseeded(a, b, 1, seed_l) * n
        cy = _t.rand_noseeded(a, b, 2, seed_l) * n
        cz = _t.rand_noseeded(a, b, 3, seed_l) * n
        r = rmin + (rmax - rmin) * _t.rand_noseeded(a, b, 4, seed_l)
        if len(drops):
            # minimum-image distance to every accepted droplet
            d = np.array([cx, cy, cz]) - acc[:, :3]
            d = d - n * np.round(d / n)
            if (np.sqrt((d ** 2).sum(axis=1)) < r + acc[:, 3] + 2.0).any():
                continue
        drops.append((cx, cy, cz, r))
        acc = np.vstack([acc, [cx, cy, cz, r]])
    return drops


def rasterize_droplets(n, drops, dtype=np.float32):
    phi = np.zeros((n, n, n), dtype=dtype)
    for (cx, cy, cz, r) in drops:
        lo = [int(np.floor(c - r)) - 1 for c in (cx, cy, cz)]
        hi = [int(np.ceil(c + r)) + 2 for c in (cx, cy, cz)]
        ax = [np.arange(lo[a], hi[a]) for a in range(3)]
        X, Y, Z = np.meshgrid(ax[0], ax[1], ax[2], indexing="ij")
        inside = (X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2 <= r * r
        xi, yi, zi = X[inside] % n, Y[inside] % n, Z[inside] % n
        phi[zi, yi, xi] = 1
    return phi


def emulsion(n=512, ndrop=200, rmin=8.0, rmax=16.0, sigma=0.01, beta=0.7, tau=0.51,
             lin_force=2.8e-5, noise=1e-2):
    """Emulsion in linearly forced homogeneous turbulence, two components,
    periodic n^3: ndrop non-overlapping droplets with radii U[rmin, rmax],
    linear forcing u_eq += A u (velocity shift, as the reference's forcing
    enters the two-component equilibrium, lb_kernels_module.f90:769-771),
    seeded velocity noise.  A = 2.8e-5 with tau = 0.51 targets Re_lambda ~ 60
    for the linear-forcing steady state (integral scale ~0.19 n).
    """
    drops = place_droplets(n, ndrop, rmin, rmax)
    phi = rasterize_droplets(n, drops)
    z = np.zeros((n, n, n), dtype=np.float32)
    vol = float(phi.sum()) / n ** 3
    return Case("emulsion", (n, n, n),
                dict(ncomp=2, tau=tau, sigma=sigma, sharp_c=3.0 * beta, lin_force=lin_force,
                     periodic=(1, 1, 1)),
                dict(rho=np.ones_like(z), u=z, v=z, w=z, phi=phi),
                noise=[("u", noise, 11), ("v", noise, 22), ("w", noise, 33)],
                notes=f"{len(drops)} droplets, volume fraction {vol:.4f}")


# ---------------------------------------------------------------------------
# newBCS turbulent jet (TSLB_newBCS/recursive_TSbcs_TSLB3.f90,
# TSLB_newBCS/boundary_cds_3D_module.f90:1233-1292)
# ---------------------------------------------------------------------------
def _inflow_disk(nx, ny, radius):
    j, i = np.meshgrid(np.arange(1, ny + 1), np.arange(1, nx + 1), indexing="ij")
    return (i - nx / 2.0) ** 2 + (j - ny / 2.0) ** 2 <= radius ** 2


def jet_inflow_initial(nx, ny, uwall=0.05, radius=10.0):
    """Initial inflow w of plane k = 1, shape (ny, nx): inside the disk
    uwall + 0.02 sqrt(-2 ln r_y) cos(2 pi r_x) with r_x, r_y the generator at
    (i, j, 1, 524) and (i, j, 1, 1732) (recursive_TSbcs_TSLB3.f90:108-125), 0
    outside."""
    rx = rand_noseeded_grid(nx, ny, 1, 524)[0]
    ry = rand_noseeded_grid(nx, ny, 1, 1732)[0]
    two_pi = np.float32(2.0 * np.pi)
    ry = np.maximum(ry, np.float32(1e-30))
    w = np.float32(uwall) + np.float32(0.02) * np.sqrt(np.float32(-2.0) * np.log(ry)) * np.cos(two_pi * rx)
    return np.where(_inflow_disk(nx, ny, radius), w, np.float32(0.0)).astype(np.float32)


def jet_inflow_redraw(w_prev, step, uwall=0.05, radius=10.0):
    """The inflow plane after the redraw of `step` (done when step % 10 == 0,
    boundary_cds_3D_module.f90:1237-1259): inside the disk
    uwall + 0.004 sqrt(-2 ln r) cos(2 pi r) with the same r = generator at
    (i, j, 1, step) twice, as the reference; outside, w_prev is kept."""
    ny, nx = w_prev.shape
    r = rand_noseeded_grid(nx, ny, 1, step)[0]
    rl = np.maximum(r, np.float32(1e-30))
    two_pi = np.float32(2.0 * np.pi)
    w = np.float32(uwall) + np.float32(0.004) * np.sqrt(np.float32(-2.0) * np.log(rl)) * np.cos(two_pi * r)
    return np.where(_inflow_disk(nx, ny, radius), w, w_prev).astype(np.float32)


def jet(nx=64, ny=64, nz=64, tau=0.6, uwall=0.05, ncomp=1, sigma=0.0, beta=0.7):
    """newBCS jet: x, y periodic; planes k = 1 and k = nz carry the moment-space
    inlet/outlet BC; u = v = w = 0, rho = 1 except w of plane k = 1, the
    initial inflow disk (radius 10).  Two components add phi = 1 on the disk of
    plane k = 1 (the inflow phase; no reference counterpart).  The caller
    applies the BC with `jet_inflow_initial` and redraws with
    `jet_inflow_redraw` (Lattice.set_zplane_bc)."""
    z = np.zeros((nz, ny, nx), dtype=np.float64)
    w = z.copy()
    w_in = jet_inflow_initial(nx, ny, uwall)
    w[0] = w_in
    phi = z.copy()
    kw = dict(ncomp=ncomp, tau=tau, periodic=(1, 1, 0))
    if ncomp == 2:
        phi[0] = _inflow_disk(nx, ny, 10.0)
        kw.update(sigma=sigma, sharp_c=3.0 * beta)
    return Case("jet", (nz, ny, nx), kw, dict(rho=np.ones_like(z), u=z, v=z.copy(), w=w, phi=phi),
                prepush_phi=ncomp == 2, notes=f"jet inflow uwall={uwall}")
