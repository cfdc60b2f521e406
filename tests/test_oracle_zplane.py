"""Pins of the oracle's moment-space z-plane BC (bcs_TSLB_only_z_turbojet,
TSLB_newBCS/boundary_cds_3D_module.f90:1233-1292) and of the newBCS jet's
inflow helpers (configs.jet_inflow_initial / jet_inflow_redraw).

The reference has no tests for this path.  The BC is pinned by
  * the closed-form steady state it must reach: with a uniform inflow W on
    plane k = 1 and the zero-gradient outlet on k = nz, a periodic duct
    settles to plug flow, w = W and a uniform mass flux rho w at every plane
    (a wrong sign, a wrong copy plane or a nonzero tangential velocity on the
    inlet leaves a different profile);
  * uniform plug flow being an exact fixed point of step + BC;
  * the bit-for-bit agreement of the population-push and moment-space pull
    forms with the BC and inflow redraws (the pull form is what the GPU runs);
  * the inflow helpers against the oracle's scalar generator, which is itself
    bit-exact against the reference-built table (test_oracle_pins.py).
"""
import numpy as np
import pytest

import pyoracle as po
from entropicthreadsafe_amd import configs


def _uniform(nx, ny, nz, rho=1.0, w=0.0):
    z = np.zeros((nz, ny, nx))
    return z + rho, z.copy(), z.copy(), z + w


def test_plug_flow_is_steady_state():
    nx, ny, nz, W = 4, 4, 48, 0.02
    o = po.Oracle(nx, ny, nz, tau=0.8, periodic=(1, 1, 0))
    rho, u, v, w = _uniform(nx, ny, nz)
    o.init(rho, u, v, w, None, False)
    o.set_zplane_bc(np.full((ny, nx), W))
    o.step(1500)
    m = o.moments()
    wm = m[3].mean(axis=(1, 2))
    flux = (m[0] * m[3]).mean(axis=(1, 2))
    # plug flow: w = W at every plane, uniform flux, no transverse flow
    assert np.abs(wm / W - 1).max() < 1e-4, wm
    assert (flux.max() - flux.min()) / flux.mean() < 1e-4
    assert np.abs(m[1]).max() < 1e-12 and np.abs(m[2]).max() < 1e-12
    # inlet plane: the inflow, zero tangential velocity; outlet: a copy
    assert np.array_equal(m[3][0], np.full((ny, nx), W))
    for q in range(10):
        assert np.array_equal(m[q][nz - 1], m[q][nz - 2])


def test_plug_flow_is_a_fixed_point():
    nx, ny, nz, W = 6, 5, 12, 0.03
    o = po.Oracle(nx, ny, nz, tau=0.6, periodic=(1, 1, 0))
    rho, u, v, w = _uniform(nx, ny, nz, rho=1.01, w=W)
    o.init(rho, u, v, w, None, False)
    o.set_zplane_bc(np.full((ny, nx), W))
    o.step(40)
    m = o.moments()
    assert np.abs(m[0] - 1.01).max() < 1e-13
    assert np.abs(m[3] - W).max() < 1e-13
    assert np.abs(m[[1, 2]]).max() < 1e-15
    assert np.abs(m[4:10]).max() < 1e-13


@pytest.mark.parametrize("ncomp", [1, 2])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_push_and_pull_agree_with_zplane_bc(ncomp, dtype):
    nx, ny, nz = 24, 22, 10
    case = configs.jet(nx, ny, nz, tau=0.7, ncomp=ncomp, sigma=0.01)
    f = case.host_fields()
    rng = np.random.default_rng(11)
    f["u"] = f["u"] + 1e-3 * rng.standard_normal(f["u"].shape)
    w_in = configs.jet_inflow_initial(nx, ny)
    phi_in = f["phi"][0] if ncomp == 2 else None
    kw = case.lattice
    out = []
    for pull in (False, True):
        o = po.Oracle(nx, ny, nz, ncomp=ncomp, tau=kw["tau"], sigma=kw.get("sigma", 0.0),
                      sharp_c=kw.get("sharp_c", 0.0), periodic=kw["periodic"], dtype=dtype)
        o.init(*[f[q].astype(dtype) for q in ("rho", "u", "v", "w", "phi")], ncomp == 2)
        o.set_zplane_bc(w_in, phi_in)
        wcur = w_in
        for s in range(1, 23):
            if s % 10 == 0:
                wcur = configs.jet_inflow_redraw(wcur, s)
                o.set_zplane_bc(wcur, phi_in)
            (o.pull_step if pull else o.step)(1)
        out.append(o.moments())
        o.close()
    assert np.isfinite(out[0]).all()
    assert np.array_equal(out[0], out[1])


def test_inflow_helpers_match_the_scalar_generator():
    nx, ny = 40, 36
    w0 = configs.jet_inflow_initial(nx, ny, uwall=0.05)
    w1 = configs.jet_inflow_redraw(w0, 30, uwall=0.05)
    inside = 0
    for j in range(1, ny + 1):
        for i in range(1, nx + 1):
            disk = (i - nx / 2.0) ** 2 + (j - ny / 2.0) ** 2 <= 100.0
            if not disk:
                assert w0[j - 1, i - 1] == 0 and w1[j - 1, i - 1] == 0
                continue
            inside += 1
            rx, ry = po.rand_noseeded(i, j, 1, 524), po.rand_noseeded(i, j, 1, 1732)
            e0 = 0.05 + 0.02 * np.sqrt(-2 * np.log(max(ry, 1e-30))) * np.cos(2 * np.pi * rx)
            r = po.rand_noseeded(i, j, 1, 30)
            e1 = 0.05 + 0.004 * np.sqrt(-2 * np.log(max(r, 1e-30))) * np.cos(2 * np.pi * r)
            assert abs(w0[j - 1, i - 1] - e0) < 2e-6
            assert abs(w1[j - 1, i - 1] - e1) < 2e-6
    assert inside == 317  # lattice points of the radius-10 disk at this centre
