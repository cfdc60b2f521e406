"""Oracle pins for the turbulent-jet boundary (bcs_2c_turbulent_jet,
boundary_cds_3D_module.f90:862-1049): regularized velocity inlet on the k=1
wall plane, outlet on k=nz, g injected by the inlet and outlet only.

The reference writes each boundary population as a closed-form polynomial in
the destination's rho, Pi and u and the wall node's velocity.  The checks below
type eight of those polynomials from the reference text and compare them with
the oracle's jet values, then run the jet in push form (the reference's
structure) and in the moment-space pull form (the GPU's structure) and require
bit-identical moments.
"""
import numpy as np
import pytest

import pyoracle as po

P1, P2, P3 = 2.0 / 27.0, 1.0 / 54.0, 1.0 / 216.0


def ref_inlet(l, om, r, P, uy, ux):
    """boundary_cds_3D_module.f90:932-983 (inlet, k=1); r = rho(y), P = Pi(y),
    (uy,vy,wy) = u(y) at the destination, (u,v,w) = u(i,j,1) of the wall node."""
    pxx, pyy, pzz, pxy, pxz, pyz = P
    u, v, w = ux
    uy_, vy, wy = uy
    if l == 5:  # :933-936
        feq = (r * (2 + 6 * w * (1 + w) - 3 * u**2 * (1 + 3 * w) - 3 * v**2 * (1 + 3 * w))) / 27.
        fneq1 = (-3 * (pxx + pyy - 2 * pzz + 6 * pxz * uy_ + 6 * pyz * vy + 3 * pxx * wy + 3 * pyy * wy)) / 2.
        return feq + (1 - om) * P1 * fneq1
    if l == 15:  # :940-942
        feq = (r * (2 + 6 * w * (1 + w) + 6 * u**2 * (1 + 3 * w) - 3 * v**2 * (1 + 3 * w)
                    + 3 * u * (2 - 3 * v**2 + 6 * w * (1 + w)))) / 108.
        fneq1 = (3 * (-pyy + 2 * pzz - 3 * pyy * uy_ + 6 * pzz * uy_ - 6 * pxy * vy - 6 * pyz * vy
                      - 3 * pyy * wy + 6 * pxz * (1 + 2 * uy_ + 2 * wy) + pxx * (2 + 6 * wy))) / 2.
        return feq + (1 - om) * P2 * fneq1
    if l == 14:  # :957-959
        feq = (r * (2 + u**2 * (-3 + 9 * v - 9 * w) + 6 * w * (1 + w) + 6 * v**2 * (1 + 3 * w)
                    - 6 * v * (1 + 3 * w * (1 + w)))) / 108.
        fneq1 = (3 * pxx * (-1 + 3 * vy - 3 * wy)) / 2. + 3 * (
            pyy + pzz + 3 * pxy * uy_ - 3 * pxz * uy_ - 3 * pzz * vy + pyz * (-3 + 6 * vy - 6 * wy)
            + 3 * pyy * wy)
        return feq + (1 - om) * P2 * fneq1
    if l == 19:  # :962-965
        feq = (r * (1 + 3 * w * (1 + w) + v**2 * (3 + 9 * w) + u**2 * (3 + 9 * v + 9 * w)
                    + v * (3 + 9 * w * (1 + w)) + 3 * u * (1 + 3 * w + 3 * (v + v**2 + 3 * v * w + w**2)))) / 216.
        fneq1 = 3 * (pxx + (pyy + 3 * pyz + pzz) * (1 + 3 * uy_) + pxy * (3 + 6 * uy_) + pxz * (3 + 6 * uy_)
                     + 3 * pxx * vy + 6 * pxy * vy + 9 * pxz * vy + 6 * pyz * vy + 3 * pzz * vy
                     + 3 * (pxx + 3 * pxy + 2 * pxz + pyy + 2 * pyz) * wy)
        return feq + (1 - om) * fneq1 * P3
    if l == 23:  # :974-977
        feq = (r * (1 + 3 * w * (1 + w) + v**2 * (3 + 9 * w) + u**2 * (3 - 9 * v + 9 * w)
                    - 3 * v * (1 + 3 * w * (1 + w))
                    - 3 * u * (1 + 3 * v**2 + 3 * w * (1 + w) - 3 * v * (1 + 3 * w)))) / 216.
        fneq1 = 3 * (pxx + 3 * pxy - 3 * pxz + pyy - 3 * pyz + pzz
                     - 3 * (2 * pxy - 2 * pxz + pyy - 3 * pyz + pzz) * uy_
                     - 3 * pxx * vy - 6 * pxy * vy + 9 * pxz * vy + 6 * pyz * vy - 3 * pzz * vy
                     + 3 * (pxx + 3 * pxy - 2 * pxz + pyy - 2 * pyz) * wy)
        return feq + (1 - om) * fneq1 * P3
    raise KeyError(l)


def ref_outlet(l, om, r, P, uy, ux):
    """boundary_cds_3D_module.f90:991-1041 (outlet, k=nz); (u,v,w) = u(i,j,nz)
    with w(i,j,nz) = w(i,j,nz-1) already copied."""
    pxx, pyy, pzz, pxy, pxz, pyz = P
    u, v, w = ux
    uy_, vy, wy = uy
    if l == 6:  # :993-995
        feq = (r * (2 + 6 * (-1 + w) * w + u**2 * (-3 + 9 * w) + v**2 * (-3 + 9 * w))) / 27.
        fneq1 = (3 * (-pyy + 2 * pzz + 6 * pxz * uy_ + 6 * pyz * vy + 3 * pyy * wy + pxx * (-1 + 3 * wy))) / 2.
        return feq + (1 - om) * fneq1 * P1
    if l == 18:  # :1003-1005
        feq = (r * (2 + u**2 * (6 - 18 * w) + 6 * (-1 + w) * w + v**2 * (-3 + 9 * w)
                    + 3 * u * (2 - 3 * v**2 + 6 * (-1 + w) * w))) / 108.
        fneq1 = (-3 * (pyy - 2 * pzz + 3 * pyy * uy_ - 6 * pzz * uy_ + 6 * pxy * vy - 6 * pyz * vy
                       + 6 * pxz * (1 + 2 * uy_ - 2 * wy) - 3 * pyy * wy + pxx * (-2 + 6 * wy))) / 2.
        return feq + (1 - om) * fneq1 * P2
    if l == 13:  # :1014-1016
        feq = (r * (2 - 6 * w + u**2 * (-3 - 9 * v + 9 * w)
                    + 6 * (v + v**2 * (1 - 3 * w) + 3 * v * (-1 + w) * w + w**2))) / 108.
        fneq1 = (-3 * (pxx - 2 * pyy + 6 * pyz - 2 * pzz + 6 * pxy * uy_ - 6 * pxz * uy_ + 3 * pxx * vy
                       + 12 * pyz * vy - 6 * pzz * vy - 3 * pxx * wy + 6 * pyy * wy - 12 * pyz * wy)) / 2.
        return feq + (1 - om) * P2 * fneq1
    if l == 24:  # :1032-1035
        feq = (r * (1 + 3 * v - 3 * w + 3 * (u + u**2 + 3 * u * v + 3 * u**2 * v + v**2 + 3 * u * v**2
                                             - 3 * (u + u**2 + v + 3 * u * v + v**2) * w
                                             + (1 + 3 * u + 3 * v) * w**2))) / 216.
        fneq1 = 3 * (pxx - 3 * pxz * (1 + 2 * uy_) + (pyy - 3 * pyz + pzz) * (1 + 3 * uy_) + pxy * (3 + 6 * uy_)
                     + 3 * pxx * vy + 6 * pxy * vy - 9 * pxz * vy - 6 * pyz * vy + 3 * pzz * vy
                     - 3 * (pxx + 3 * pxy - 2 * pxz + pyy - 2 * pyz) * wy)
        return feq + (1 - om) * P3 * fneq1
    raise KeyError(l)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_jet_values_match_reference_polynomials(seed):
    rng = np.random.default_rng(seed)
    om = 1.0 / 0.53
    r = 1.0 + 0.05 * rng.standard_normal()
    P = 1e-2 * rng.standard_normal(6)
    uy = 0.05 * rng.standard_normal(3)
    ux = 0.05 * rng.standard_normal(3)
    got = po.jet_values(om, r, P, uy, ux)
    for l in (5, 15, 14, 19, 23):
        assert got[l] == pytest.approx(ref_inlet(l, om, r, P, uy, ux), rel=1e-13, abs=1e-15), l
    for l in (6, 18, 13, 24):
        assert got[l] == pytest.approx(ref_outlet(l, om, r, P, uy, ux), rel=1e-13, abs=1e-15), l


def test_jet_value_roles_are_not_interchangeable():
    """A planted mistake: swapping the wall velocity and the destination
    velocity (or taking rho from the wall) must fail the pins above."""
    rng = np.random.default_rng(5)
    om = 1.0 / 0.53
    r = 1.02
    P = 1e-2 * rng.standard_normal(6)
    uy = 0.05 * rng.standard_normal(3)
    ux = 0.05 * rng.standard_normal(3)
    swapped = po.jet_values(om, r, P, ux, uy)
    assert abs(swapped[15] - ref_inlet(15, om, r, P, uy, ux)) > 1e-6


def jet_case(nx=10, ny=10, nz=12, ncomp=2, seed=3):
    """Reference jet geometry (recursive_2c_TSLB3D.f90:79-100): walls on all six
    faces, inflow disk on the k=1 plane with phi=1 and a noisy w."""
    rng = np.random.default_rng(seed)
    isf = np.ones((nz, ny, nx), np.int32)
    isf[0], isf[-1] = 0, 0
    isf[:, 0, :], isf[:, -1, :] = 0, 0
    isf[:, :, 0], isf[:, :, -1] = 0, 0
    rho = (1 + 1e-3 * rng.standard_normal((nz, ny, nx)))
    u = 1e-3 * rng.standard_normal((nz, ny, nx))
    v = 1e-3 * rng.standard_normal((nz, ny, nx))
    w = 1e-3 * rng.standard_normal((nz, ny, nx))
    phi = np.zeros((nz, ny, nx))
    jj, ii = np.mgrid[0:ny, 0:nx]
    disk = (ii + 1 - nx / 2) ** 2 + (jj + 1 - ny / 2) ** 2 <= 3.0 ** 2
    w[0] = np.where(disk, 0.05 + 0.004 * rng.standard_normal((ny, nx)), 0.0)
    u[0] = v[0] = 0.0
    phi[0] = disk.astype(float)
    kw = dict(ncomp=ncomp, tau=0.56, sigma=0.01 if ncomp == 2 else 0.0,
              sharp_c=2.1 if ncomp == 2 else 0.0, periodic=(0, 0, 0))
    return isf, (rho, u, v, w, phi), kw


def run_jet(form, ncomp, steps, dtype=np.float64):
    isf, fields, kw = jet_case(ncomp=ncomp)
    nz, ny, nx = isf.shape
    o = po.Oracle(nx, ny, nz, dtype=dtype, **kw)
    o.set_isfluid(isf)
    o.set_jet(True)
    o.init(*fields, prepush=True)
    for _ in range(steps):
        (o.step if form == "push" else o.pull_step)(1)
    m = o.moments()
    o.close()
    return m


@pytest.mark.parametrize("ncomp", [1, 2])
def test_jet_push_and_pull_forms_agree_bitwise(ncomp):
    a = run_jet("push", ncomp, 6)
    b = run_jet("pull", ncomp, 6)
    assert np.isfinite(a).all()
    assert np.array_equal(a, b)


def test_jet_injects_phi_and_copies_outlet_velocity():
    isf, fields, kw = jet_case(ncomp=2)
    nz, ny, nx = isf.shape
    o = po.Oracle(nx, ny, nz, **kw)
    o.set_isfluid(isf)
    o.set_jet(True)
    o.init(*fields, prepush=True)
    phi_in = o.moments()[10][1].sum()
    o.step(8)
    m = o.moments()
    o.close()
    # the inlet disk carries phi into the first fluid planes
    assert m[10][1:4].sum() > phi_in
    # outlet wall plane holds w of the plane below it (of the previous step's
    # moments; after a step the copy lags by one step, so compare with a run
    # one step shorter)
    o2 = po.Oracle(nx, ny, nz, **kw)
    o2.set_isfluid(isf)
    o2.set_jet(True)
    o2.init(*fields, prepush=True)
    o2.step(7)
    m7 = o2.moments()
    o2.close()
    assert np.array_equal(m[3][nz - 1], m7[3][nz - 2])
