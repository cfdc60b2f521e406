"""Closed-form pins of the oracle's interface terms (recursive_2c_TSLB3D.f90):
the gated curvature kappa = -div(n) (:614-631), the sharpened D3Q7 phase
field (:808-831) and the CSF pressure jump (:632-636 with the step).

  * kappa of a smooth sphere is 2/r and of a cylinder 1/r at every node past
    the gate phi(1-phi)|n|^2 > 0.2 (a dropped direction, a wrong weight or a
    sign slip in the stencil or the divergence moves these by O(1));
  * a flat interface at rest relaxes to the closed-form steady profile of the
    D3Q7 update with sharpening: with u = 0 the update is
    phi'(x) = 3/4 phi(x) + (phi(x-1) + phi(x+1))/8 + (s(x-1) - s(x+1))/8,
    s = sharp_c phi (1 - phi) n_x, whose steady state obeys
    phi'' = 2 s', i.e. phi = (1 - tanh(sharp_c (x - x0)))/2 in the continuum;
    the coefficient is pinned by a planted halved sharp_c, which the check
    rejects;
  * the pressure jump of a resting droplet is linear in sigma (time averaged:
    the set of nodes past the gate changes from step to step, so a single
    sample moves by ~15 %), and of the order of 2 sigma / R.
"""
import numpy as np

import pyoracle as po


def _set_phi(o, phi):
    n = phi.shape
    m = np.zeros((11,) + n)
    m[0] = 1.0
    m[10] = phi
    o.set_moments(m)


def test_curvature_of_a_sphere_and_a_cylinder():
    n, R, W = 64, 16.0, 2.0
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    c = n / 2 - 0.5
    for shape in ("sphere", "cylinder"):
        if shape == "sphere":
            r = np.sqrt((i - c) ** 2 + (j - c) ** 2 + (k - c) ** 2)
            exact = 2.0 / np.maximum(r, 1e-9)
        else:
            r = np.sqrt((i - c) ** 2 + (j - c) ** 2)
            exact = 1.0 / np.maximum(r, 1e-9)
        o = po.Oracle(n, n, n, ncomp=2, tau=1.0, sigma=0.01, sharp_c=2.1)
        _set_phi(o, 0.5 * (1 + np.tanh((R - r) / W)))
        kap = o.curvature()
        o.close()
        gated = kap != 0
        assert gated.sum() > 1000
        ratio = kap[gated] / exact[gated]
        print(shape, gated.sum(), "kappa / exact: min %.5f max %.5f" % (ratio.min(), ratio.max()))
        assert np.abs(ratio - 1).max() < 1e-2
        # the gate holds the shell around r = R (phi(1-phi) > 0.2 there)
        assert np.abs(r[gated] - R).max() < W
        mean_exact = 2.0 / R if shape == "sphere" else 1.0 / R
        assert abs(kap[gated].mean() / mean_exact - 1) < 1e-2


def _flat_profile(sharp_c, steps=3000, nx=64):
    o = po.Oracle(nx, 4, 4, ncomp=2, tau=1.0, sigma=0.0, sharp_c=sharp_c)
    x = np.arange(nx)
    phi = ((x >= 16) & (x < 48)).astype(float)[None, None, :] * np.ones((4, 4, 1))
    z = np.zeros_like(phi)
    o.init(np.ones_like(phi), z, z, z, phi, False)
    o.step(steps)
    m = o.moments()
    o.close()
    return m


def _profile_deviation(p, sharp_c):
    x = np.arange(p.size)
    i1 = np.where((p[:-1] < 0.5) & (p[1:] >= 0.5))[0][0]
    i2 = np.where((p[:-1] >= 0.5) & (p[1:] < 0.5))[0][0]
    x1 = i1 + (0.5 - p[i1]) / (p[i1 + 1] - p[i1])
    x2 = i2 + (p[i2] - 0.5) / (p[i2] - p[i2 + 1])
    model = 0.5 * (np.tanh(sharp_c * (x - x1)) - np.tanh(sharp_c * (x - x2)))
    return np.abs(p - model).max(), x1, x2


def test_flat_interface_steady_profile():
    for sharp_c in (0.3, 0.5):
        m = _flat_profile(sharp_c)
        p = m[10][0, 0, :]
        dev, x1, x2 = _profile_deviation(p, sharp_c)
        print("sharp_c", sharp_c, "interfaces at", x1, x2, "max deviation from the tanh profile", dev)
        assert dev < 1e-2
        # at rest, no velocity appears (sigma = 0) and phi is conserved
        assert np.abs(m[1:4]).max() == 0.0
        assert abs(p.sum() - 32.0) < 1e-10
        # a planted halved coefficient gives the profile of sharp_c / 2
        planted = _flat_profile(0.5 * sharp_c)[10][0, 0, :]
        assert _profile_deviation(planted, sharp_c)[0] > 0.1


def _mean_jump(n, R, sigma, t0=500, t1=1500, every=5, tau=1.5, sharp_c=2.1):
    o = po.Oracle(n, n, n, ncomp=2, tau=tau, sigma=sigma, sharp_c=sharp_c)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    c = n / 2 - 0.5
    r = np.sqrt((i - c) ** 2 + (j - c) ** 2 + (k - c) ** 2)
    phi = 0.5 * (1 + np.tanh(sharp_c * (R - r)))
    z = np.zeros_like(phi)
    o.init(np.ones_like(phi), z, z, z, phi)
    o.step(t0)
    ins, outs = r < R - 3, r > R + 4
    js = []
    for _ in range((t1 - t0) // every):
        o.step(every)
        m = o.moments()
        js.append((m[0][ins].mean() - m[0][outs].mean()) / 3.0)
    o.close()
    return float(np.mean(js))


def test_laplace_jump_is_linear_in_sigma():
    n, R = 40, 12.0
    d2 = _mean_jump(n, R, 0.02)
    d1 = _mean_jump(n, R, 0.01)
    print("time-averaged jump / (2 sigma / R):", d2 / (2 * 0.02 / R), d1 / (2 * 0.01 / R))
    assert abs(d2 / (2 * d1) - 1) < 0.05
    # the CSF shift acts on the gated band only, which gives an effective
    # surface tension below sigma for this sharp interface
    assert 0.5 < d2 / (2 * 0.02 / R) < 1.2
