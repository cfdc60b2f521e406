"""Oracle pins for the colour-gradient model (ncomp 3).

The formula source (turbulent_benchmark/couette/recursive_colorTSLB3D.f90:
719-959) does not compile and has defects, so the oracle restates the
textbook model with the file's coefficients; these tests pin that restatement
against what the mathematics fixes: conservation of mass and momentum by the
perturbation, its interfacial stress, conservation of each component's mass
by the recolouring, reduction to the one-component step without interface,
push/pull equivalence, and the Laplace law.
"""
import numpy as np
import pytest

import pyoracle as po
from test_oracle_pins import CX, CY, CZ, W

C = np.stack([CX, CY, CZ], axis=1).astype(float)


def _random_node(seed):
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.02 * rng.standard_normal()
    u = 0.03 * rng.standard_normal(3)
    P = 1e-3 * rng.standard_normal(6)
    psi = 0.6 * rng.uniform(-1, 1)
    g = 0.2 * rng.standard_normal(3)
    return rho, u, P, psi, g


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_perturbation_conserves_mass_momentum_and_gives_the_interfacial_stress(seed):
    rho, u, P, psi, g = _random_node(seed)
    om, sigma = 1.0 / 0.7, 0.03
    fo, _ = po.colour_values(om, om, sigma, 0.0, rho, u, P, psi, g)
    f0, _ = po.colour_values(om, om, 0.0, 0.0, rho, u, P, psi, g)
    d = fo - f0                      # the perturbation Omega_l
    assert abs(d.sum()) < 1e-15
    assert np.abs(C.T @ d).max() < 1e-15
    # second moment: (A/(9|g|)) (g g - |g|^2 I), A = (9/2) sigma omega
    A = 4.5 * sigma * om
    mg = np.linalg.norm(g)
    M2 = (C * d[:, None]).T @ C
    exp = A / (9 * mg) * (np.outer(g, g) - mg ** 2 * np.eye(3))
    assert np.abs(M2 - exp).max() < 1e-14


def test_planted_linear_perturbation_breaks_momentum():
    """The formula source's perturbation is linear in c.g; that form would
    move momentum, which the check above rejects."""
    rho, u, P, psi, g = _random_node(7)
    A, mg = 0.1, np.linalg.norm(g)
    k = np.array([1.0] + [27.0] * 6 + [108.0] * 12 + [432.0] * 8)
    cg = C @ g
    lin = A * (cg - mg ** 2) / (k * mg)
    lin[0] = 5 * A * mg / 27
    assert np.abs(C.T @ lin).max() > 1e-3


@pytest.mark.parametrize("seed", [0, 1])
def test_recolouring_conserves_component_mass(seed):
    rho, u, P, psi, g = _random_node(seed)
    om = 1.0 / 0.7
    fo, fa = po.colour_values(om, 1.0 / 0.9, 0.02, 0.7, rho, u, P, psi, g)
    ca = 0.5 * (1 + psi)
    assert abs(fa.sum() - ca * fo.sum()) < 1e-15
    # the A populations lean towards +grad(psi): positive A flux along g
    assert (C.T @ (fa - ca * fo)) @ g > 0
    # Latva-Kokko term: beta cA cB w rho cos(theta)
    fb = fo - fa
    nrm = np.sqrt((C ** 2).sum(1))
    cos = np.where(nrm > 0, (C @ g) / np.where(nrm > 0, nrm, 1) / np.linalg.norm(g), 0)
    exp = 0.7 * ca * (1 - ca) * W * rho * cos
    assert np.abs((fa - ca * fo) - exp).max() < 1e-15
    assert np.abs((fb - (1 - ca) * fo) + exp).max() < 1e-15


def _colour_case(n=12, seed=4):
    rng = np.random.default_rng(seed)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    r = np.sqrt((i - n / 2 + 0.5) ** 2 + (j - n / 2 + 0.5) ** 2 + (k - n / 2 + 0.5) ** 2)
    psi = np.tanh(0.8 * (n / 4 - r))
    rho = 1 + 1e-3 * rng.standard_normal(psi.shape)
    u = 1e-3 * rng.standard_normal(psi.shape)
    v = 1e-3 * rng.standard_normal(psi.shape)
    w = 1e-3 * rng.standard_normal(psi.shape)
    return rho, u, v, w, psi


@pytest.mark.parametrize("walls", [False, True])
def test_colour_push_and_pull_forms_agree_bitwise(walls):
    n = 12
    fields = _colour_case(n)
    res = []
    for form in ("push", "pull"):
        o = po.Oracle(n, n, n, ncomp=3, tau=0.7, sigma=0.02, force=(1e-6, 0, 0),
                      periodic=(1, 0, 1) if walls else (1, 1, 1))
        o.set_colour(0.9, 0.7)
        if walls:
            isf = np.ones((n, n, n), np.int32)
            isf[5:7, 2:4, 3:9] = 0
            o.set_isfluid(isf)
        o.init(*fields)
        for _ in range(5):
            (o.step if form == "push" else o.pull_step)(1)
        res.append(o.moments())
        o.close()
    assert np.isfinite(res[0]).all()
    assert np.array_equal(res[0], res[1])


def test_colour_reduces_to_one_component_without_interface():
    """psi uniform: no gradient, no perturbation or recolouring, one
    viscosity; the total moments follow the one-component step."""
    n = 12
    rho, u, v, w, _ = _colour_case(n)
    psi = np.full_like(rho, 0.4)
    a = po.Oracle(n, n, n, ncomp=3, tau=0.7, sigma=0.02, force=(1e-6, 0, 0))
    a.set_colour(0.7, 0.7)
    a.init(rho, u, v, w, psi)
    a.step(8)
    b = po.Oracle(n, n, n, ncomp=1, tau=0.7, force=(1e-6, 0, 0))
    b.init(rho, u, v, w)
    b.step(8)
    ma, mb = a.moments(), b.moments()
    assert np.abs(ma[:10] - mb[:10]).max() < 1e-13
    assert np.abs(ma[10] - 0.4).max() < 1e-12


def test_component_masses_are_conserved():
    n = 12
    fields = _colour_case(n)
    o = po.Oracle(n, n, n, ncomp=3, tau=0.7, sigma=0.02)
    o.set_colour(0.9, 0.7)
    o.init(*fields)
    r0 = o.reduce()
    o.step(50)
    r1 = o.reduce()
    o.close()
    assert abs(r1[0] - r0[0]) < 1e-11 * r0[0]
    assert abs(r1[1] - r0[1]) < 1e-11 * r0[1]


def test_colour_laplace_law():
    """Static droplet: rho_in - rho_out = 2 sigma / (cs^2 R)."""
    n, R, sigma = 32, 8.0, 0.02
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    r = np.sqrt((i - 15.5) ** 2 + (j - 15.5) ** 2 + (k - 15.5) ** 2)
    psi = np.tanh(R - r)
    z = np.zeros_like(psi)
    o = po.Oracle(n, n, n, ncomp=3, tau=1.0, sigma=sigma)
    o.set_colour(1.0, 0.7)
    o.init(np.ones_like(psi), z, z, z, psi)
    o.step(500)
    m = o.moments()
    o.close()
    dp = (m[0][r < R - 3].mean() - m[0][r > R + 5].mean()) / 3.0
    print("colour Laplace", dp, 2 * sigma / R)
    assert dp > 0
    assert abs(dp / (2 * sigma / R) - 1) < 0.1
    # the interface stays sharp: psi still spans most of [-1, 1]
    assert m[10].max() > 0.9 and m[10].min() < -0.9
