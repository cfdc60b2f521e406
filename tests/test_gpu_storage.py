"""Reduced-precision moment storage (FP16 / BF16 / FX16, FP32 compute) against
the FP64 oracle.

The state is stored with a shifted density (rho - 1) and, for FP16, scaled
rho-1 (x64) and Pi^neq (x4096); FX16 stores 16-bit fixed point over a half
range per field (rho-1 1/4, u 0.375, Pi^neq 1/8, phi 1.5); every kernel
computes in FP32.  Errors are the
per-field max relative errors of tests/parity.py (RMS errors are printed too).

Each step rounds every stored field once, and Pi^neq at tau ~ 0.51 decays by
only |1 - omega| ~ 0.93-0.96 per step, so rounding accumulates over ~15-25
steps.  Measured on the MI355X (max over TGV 32^3 x 100 steps and channel x
300 steps; rho, u, Pi): FP16 1.2e-4, 1.6e-3, 1.4e-2; BF16 (8 significant
bits) 5.2e-4, 2.9e-2, 0.12; FX16 3.4e-5, 7.6e-4, 5.7e-3.  The bounds are 1.5x
those, so a broken codec or scale fails them.
"""
import numpy as np
import pytest

from entropicthreadsafe_amd import configs
from entropicthreadsafe_amd.tslb import Lattice
from parity import field_errors, field_rms_errors
from test_gpu_parity import run_oracle

pytestmark = pytest.mark.gpu

# (rho, u, Pi) bounds per storage: 1.5x the larger of the errors measured on
# the MI355X in the Taylor-Green and channel cases below
TOL = {"fp16": (1.6e-4, 3.7e-3, 2.6e-2), "bf16": (1.3e-3, 2.7e-2, 0.2), "fx16": (5.3e-5, 1.3e-3, 8.3e-3)}
# the same at the C3 size, where Pi^neq carries the noise of the stored
# velocity's rounding (test_c3_channel_storage_error); FX16 meets the 1e-3
# target for rho and u (measured 5.9e-5 and 4.5e-4)
TOL_C3 = {"fp16": (3.0e-4, 4.5e-3, 0.33), "bf16": (2.7e-3, 3.4e-2, 2.4), "fx16": (8.8e-5, 6.8e-4, 9.1e-2)}
STORAGES = ["fp16", "bf16", "fx16"]


def check_bounds(err, storage):
    t_rho, t_u, t_pi = TOL[storage]
    assert err["rho"] <= t_rho, ("rho", err["rho"])
    for k in ("u", "v", "w"):
        assert err[k] <= t_u, (k, err[k])
    for k in ("pxx", "pyy", "pzz", "pxy", "pxz", "pyz"):
        assert err[k] <= t_pi, (k, err[k])


def run_storage(kw, fields, steps, storage):
    nz, ny, nx = fields["rho"].shape
    with Lattice(nx, ny, nz, storage=storage, **kw) as lat:
        for name in ("rho", "u", "v", "w", "phi"):
            lat.upload_field(name, fields[name])
        lat.init_equilibrium(True)
        lat.step(steps)
        return lat.download()


@pytest.mark.parametrize("storage", STORAGES)
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
    # digits of rho-1, and of u; FX16 is within half a step, range / 65534
    if storage == "fx16":
        assert np.abs(back - f).max() <= 0.25 / 65534 * 1.001
        assert np.abs(ub - u).max() <= 0.375 / 65534 * 1.001
        return
    rel = 1e-3 if storage == "fp16" else 8e-3
    assert np.abs(back - f).max() <= rel * np.abs(f - 1).max()
    assert np.abs(ub - u).max() <= rel * np.abs(u).max()


@pytest.mark.parametrize("storage", STORAGES)
def test_tgv_reduced_storage_vs_fp64(storage):
    case = configs.tgv(n=32, u0=0.05, re=400.0)
    f = case.host_fields()
    got = run_storage(case.lattice, f, 100, storage)
    ref = run_oracle(case.lattice, f, 100)
    err = field_errors(got, ref)
    print(storage, "max", {k: f"{v:.2e}" for k, v in err.items()})
    print(storage, "rms", {k: f"{v:.2e}" for k, v in field_rms_errors(got, ref).items()})
    check_bounds(err, storage)


@pytest.mark.parametrize("storage", STORAGES)
def test_channel_reduced_storage_vs_fp64(storage):
    case = configs.channel(nx=32, ny=24, nz=16, re_tau=10.0, u_tau=0.005, pert=0.05)
    f = case.host_fields()
    got = run_storage(case.lattice, f, 300, storage)
    ref = run_oracle(case.lattice, f, 300)
    err = field_errors(got, ref)
    print(storage, "max", {k: f"{v:.2e}" for k, v in err.items()})
    print(storage, "rms", {k: f"{v:.2e}" for k, v in field_rms_errors(got, ref).items()})
    check_bounds(err, storage)


@pytest.mark.parametrize("storage", STORAGES)
def test_slow_forcing_moves_a_uniform_flow(storage):
    """A uniform flow u0 = 0.05 under a body force of 2e-7 per step, fully
    periodic: over 2000 steps its velocity grows by 4e-4, every step far
    below half a storage step at u0 (BF16 2.4e-4, FP16 3.1e-5, FX16
    7.6e-6).  Rounded to nearest, the stored velocity never moves; the
    fused pass rounds u stochastically, so its mean grows as in FP32
    storage, within the noise of the rounding (~5 % for BF16 over these
    65 536 nodes)."""
    nx, ny, nz = 64, 32, 32
    fx, steps, u0 = 2e-7, 2000, 0.05
    shape = (nz, ny, nx)
    du = {}
    for st in ("fp32", storage):
        with Lattice(nx, ny, nz, tau=0.8, force=(fx, 0.0, 0.0), storage=st) as lat:
            lat.upload_field("rho", np.ones(shape, np.float32))
            lat.upload_field("u", np.full(shape, u0, np.float32))
            lat.upload_field("v", np.zeros(shape, np.float32))
            lat.upload_field("w", np.zeros(shape, np.float32))
            lat.init_equilibrium(False)
            a = lat.download_field("u").astype(np.float64).mean()
            lat.step(steps)
            du[st] = lat.download_field("u").astype(np.float64).mean() - a
    print(storage, "mean u growth", du[storage], "fp32", du["fp32"], "force x steps", steps * fx)
    assert abs(du["fp32"] / (steps * fx) - 1) < 0.01
    assert abs(du[storage] / du["fp32"] - 1) < 0.25


@pytest.mark.parametrize("storage", STORAGES)
def test_droplet_reduced_storage_conserves_mass(storage):
    case = configs.droplet(n=32, radius=8.0, sigma=0.02, beta=0.7, noise=1e-3)
    with Lattice(32, 32, 32, storage=storage, **case.lattice) as lat:
        case.apply(lat)
        r0 = lat.reduce()
        lat.step(200)
        r1 = lat.reduce()
        m = lat.download()
        clipped = lat.storage_clipped()
    print(storage, "mass drift", (r1[0] - r0[0]) / r0[0], "phase mass drift", (r1[1] - r0[1]) / r0[1],
          "clipped", clipped)
    # the start-up of the sharp droplet stays inside the FX16 ranges
    assert clipped == 0
    assert np.isfinite(m).all()
    # storage rounding of rho-1 and phi each step: a random walk well below 1e-4
    assert abs(r1[0] - r0[0]) < 1e-4 * r0[0]
    assert abs(r1[1] - r0[1]) < 1e-2 * r0[1]


@pytest.mark.parametrize("storage", STORAGES)
def test_c3_channel_storage_error(storage):
    """The C3 size itself (channel 512x256x256, Re_tau 180, 200 steps):
    reduced storage against the FP32 path, which tracks FP64 within 1e-5
    (tests/test_gpu_parity.py).  Prints max and RMS errors per field.

    Here Pi^neq (~nu du/dy ~ 4e-6) is of the order of the velocity gradients
    that the rounding of the stored u makes (FP16 steps 6e-5 at u ~ 0.08,
    BF16 5e-4), so Pi^neq carries that noise: the bounds on Pi are far above
    the small cases' (measured FP16: rho 2.0e-4, u 3.0e-3, Pi 0.22 max; RMS
    3.2e-5, 1.1e-3, 0.04-0.17)."""
    case = configs.channel()
    nz, ny, nx = case.shape
    out = {}
    for st in ("fp32", storage):
        with Lattice(nx, ny, nz, storage=st, **case.lattice) as lat:
            case.apply(lat)
            lat.step(200)
            out[st] = lat.download()
    ref, got = out["fp32"], out[storage]
    err = field_errors(got, ref)
    rms = field_rms_errors(got, ref)
    print(storage, "C3 max", {k: f"{v:.2e}" for k, v in err.items() if k != "phi"})
    print(storage, "C3 rms", {k: f"{v:.2e}" for k, v in rms.items() if k != "phi"})
    t_rho, t_u, t_pi = TOL_C3[storage]
    assert err["rho"] <= t_rho, ("rho", err["rho"])
    for k in ("u", "v", "w"):
        assert err[k] <= t_u, (k, err[k])
    for k in ("pxx", "pyy", "pzz", "pxy", "pxz", "pyz"):
        assert err[k] <= t_pi, (k, err[k])


def test_fx16_range_saturates_and_reencodes():
    n = 8
    u = np.linspace(-0.5, 0.5, n ** 3, dtype=np.float32).reshape(n, n, n)
    step = 2.0 ** -17  # the default step of u, v, w: range 32767 * 2^-17 ~ 0.25
    with Lattice(n, n, n, storage="fx16") as lat:
        lat.upload_field("u", u)
        got = lat.download_field("u")
        # beyond the range the field is clipped, and counted
        assert np.abs(got).max() == np.float32(32767 * step)
        assert lat.storage_clipped() == int((np.abs(u.astype(np.float64) / step) > 32767.5).sum())
        assert np.abs(got - np.clip(u, -32767 * step, 32767 * step)).max() <= step / 2
        # a wider range takes the finest power-of-two step that covers it
        # (0.75 -> 2^-15, range ~1) and re-encodes the stored field
        lat.set_storage_range("u", 0.75)
        again = lat.download_field("u")
        assert np.abs(again - got).max() <= 2.0 ** -16
        lat.upload_field("u", u)
        assert np.abs(lat.download_field("u") - u).max() <= 2.0 ** -16
    with Lattice(n, n, n, storage="fp16") as lat:
        with pytest.raises(Exception):
            lat.set_storage_range("u", 1.0)
