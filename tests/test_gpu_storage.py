"""Reduced-precision moment storage (FP16 / BF16, FP32 compute) against the
FP64 oracle.

The state is stored with a shifted density (rho - 1) and, for FP16, scaled
rho-1 (x64) and Pi^neq (x4096); every kernel computes in FP32.  Errors are the
per-field max relative errors of tests/parity.py.  Bounds:
  * FP16 storage: <= 1e-3 on every field (the precision target);
  * BF16 storage: <= 1e-2 on every field (BF16 keeps 8 significant bits: one
    rounding of u is ~2e-3 of the velocity scale, so the 1e-3 target is out
    of reach for BF16 and the measured figure is reported).
"""
import numpy as np
import pytest

from entropicthreadsafe_amd import configs
from entropicthreadsafe_amd.tslb import Lattice
from parity import field_errors
from test_gpu_parity import run_oracle

pytestmark = pytest.mark.gpu

TOL = {"fp16": 1e-3, "bf16": 1e-2}


def run_storage(kw, fields, steps, storage):
    nz, ny, nx = fields["rho"].shape
    with Lattice(nx, ny, nz, storage=storage, **kw) as lat:
        for name in ("rho", "u", "v", "w", "phi"):
            lat.upload_field(name, fields[name])
        lat.init_equilibrium(True)
        lat.step(steps)
        return lat.download()


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_storage_roundtrip(storage):
    n = 16
    rng = np.random.default_rng(11)
    f = (1 + 1e-3 * rng.standard_normal((n, n, n))).astype(np.float32)
    u = (0.05 * rng.standard_normal((n, n, n))).astype(np.float32)
    with Lattice(n, n, n, storage=storage) as lat:
        lat.upload_field("rho", f)
        lat.upload_field("u", u)
        back = lat.download_field("rho")
        ub = lat.download_field("u")
    # rho-1 is stored: the round trip keeps ~3 (FP16) or ~2 (BF16) significant
    # digits of rho-1, and of u
    rel = 1e-3 if storage == "fp16" else 8e-3
    assert np.abs(back - f).max() <= rel * np.abs(f - 1).max()
    assert np.abs(ub - u).max() <= rel * np.abs(u).max()


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_tgv_reduced_storage_vs_fp64(storage):
    case = configs.tgv(n=32, u0=0.05, re=400.0)
    f = case.host_fields()
    got = run_storage(case.lattice, f, 100, storage)
    ref = run_oracle(case.lattice, f, 100)
    err = field_errors(got, ref)
    print(storage, {k: f"{v:.2e}" for k, v in err.items()})
    for k in ("rho", "u", "v", "w", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz"):
        assert err[k] <= TOL[storage], (k, err[k])


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_channel_reduced_storage_vs_fp64(storage):
    case = configs.channel(nx=32, ny=24, nz=16, re_tau=10.0, u_tau=0.005, pert=0.05)
    f = case.host_fields()
    got = run_storage(case.lattice, f, 300, storage)
    ref = run_oracle(case.lattice, f, 300)
    err = field_errors(got, ref)
    print(storage, {k: f"{v:.2e}" for k, v in err.items()})
    for k in ("rho", "u", "v", "w", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz"):
        assert err[k] <= TOL[storage], (k, err[k])


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_droplet_reduced_storage_conserves_mass(storage):
    case = configs.droplet(n=32, radius=8.0, sigma=0.02, beta=0.7, noise=1e-3)
    with Lattice(32, 32, 32, storage=storage, **case.lattice) as lat:
        case.apply(lat)
        r0 = lat.reduce()
        lat.step(200)
        r1 = lat.reduce()
        m = lat.download()
    assert np.isfinite(m).all()
    # storage rounding of rho-1 and phi each step: a random walk well below 1e-4
    assert abs(r1[0] - r0[0]) < 1e-4 * r0[0]
    assert abs(r1[1] - r0[1]) < 1e-2 * r0[1]
