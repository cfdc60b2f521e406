"""The emulsion workload (two components, many droplets, linear forcing,
seeded velocity noise; configs.emulsion) on the GPU.

  * A reduced emulsion (64^3, 8 droplets with radii U[4, 8]) against the FP64
    oracle, pointwise over 15 steps: FP32 storage within 1e-5 per field (the
    two-component horizon of tests/test_gpu_parity.py), FP16, BF16 and FX16
    storage within the bounds below (about 1.5x the measured errors).
  * The workload at its size (512^3, 200 droplets with radii U[8, 16]) over
    200 steps in FP32, BF16 and FX16 storage: finite state, mass drift (FP64
    sums) and phase-mass drift within stated bounds.

Reference step: recursive_2c_TSLB3D.f90:344-840; linear forcing enters the
two-component equilibrium as a velocity shift (TSLB_2C-mpi/lb_kernels_module.
f90:769-771).
"""
import numpy as np
import pytest

import pyoracle as po
from entropicthreadsafe_amd import configs
from entropicthreadsafe_amd.tslb import Lattice
from parity import drho_error, field_errors, field_rms_errors

pytestmark = pytest.mark.gpu

# Reduced storage against the FP64 oracle after 15 steps: max and RMS error
# per field group (tests/parity.py scales), about 1.5x the values measured on
# the MI355X.  The max is set by the curvature gate phi(1-phi)|n|^2 > 0.2
# (recursive_2c_TSLB3D.f90:615-616): 16-bit phi moves a few interface nodes
# across it, and a node that gains or loses its CSF shift sigma*kappa*n
# (~3e-3 here, against |u| ~ 1e-2) differs by tens of percent of the velocity
# scale while the RMS stays at the storage rounding.
TOL_SMALL = {
    "fp16": dict(max=dict(rho=1.3e-3, u=0.14, p=0.28, phi=1.9e-2),
                 rms=dict(rho=1.2e-5, u=8.7e-3, p=1.2e-2, phi=1.3e-3)),
    "bf16": dict(max=dict(rho=1.8e-3, u=0.3, p=0.33, phi=0.1),
                 rms=dict(rho=4.3e-5, u=2.6e-2, p=3.5e-2, phi=8.7e-3)),
    "fx16": dict(max=dict(rho=6.8e-4, u=0.12, p=5.7e-2, phi=5.9e-3),
                 rms=dict(rho=1.1e-5, u=5.9e-3, p=7.8e-3, phi=3.2e-4)),
}
GROUPS = dict(rho=("rho",), u=("u", "v", "w"), p=("pxx", "pyy", "pzz", "pxy", "pxz", "pyz"), phi=("phi",))


def small_emulsion():
    return configs.emulsion(n=64, ndrop=8, rmin=4.0, rmax=8.0, noise=1e-2)


def _run(case, steps, storage):
    n = case.shape[0]
    f = case.host_fields()
    with Lattice(n, n, n, storage=storage, **case.lattice) as lat:
        for name in ("rho", "u", "v", "w", "phi"):
            lat.upload_field(name, f[name])
        lat.init_equilibrium(True)
        r0 = lat.reduce()
        lat.step(steps)
        r1 = lat.reduce()
        got = lat.download()
    kw = case.lattice
    o = po.Oracle(n, n, n, ncomp=2, tau=kw["tau"], sigma=kw["sigma"], sharp_c=kw["sharp_c"],
                  lin_a=kw["lin_force"], periodic=kw["periodic"])
    o.init(*[f[q].astype(np.float64) for q in ("rho", "u", "v", "w", "phi")], True)
    o.step(steps)
    ref = o.moments()
    o.close()
    return got, ref, r0, r1


def test_small_emulsion_fp32_matches_fp64_oracle():
    case = small_emulsion()
    got, ref, r0, r1 = _run(case, 15, "fp32")
    e = field_errors(got, ref)
    dr = drho_error(got, ref)
    print("emulsion 64^3 fp32 vs FP64", {k: f"{v:.2e}" for k, v in e.items()}, f"drho {dr:.2e}")
    assert np.isfinite(got).all()
    # the droplets carry interface nodes past the curvature gate
    phi = ref[10]
    assert (phi * (1 - phi) > 0.2).sum() > 500
    for k, v in e.items():
        assert v <= 1e-5, (k, v)
    assert dr <= 1e-5
    assert abs(r1[0] - r0[0]) <= 1e-9 * r0[0]
    assert abs(r1[1] - r0[1]) <= 1e-7 * r0[1]


@pytest.mark.parametrize("storage", ["fp16", "bf16", "fx16"])
def test_small_emulsion_reduced_storage(storage):
    case = small_emulsion()
    got, ref, r0, r1 = _run(case, 15, storage)
    e = field_errors(got, ref)
    rms = field_rms_errors(got, ref)
    print(f"emulsion 64^3 {storage} vs FP64 max", {k: f"{v:.2e}" for k, v in e.items()})
    print(f"emulsion 64^3 {storage} vs FP64 rms", {k: f"{v:.2e}" for k, v in rms.items()})
    print(f"mass drift {(r1[0] - r0[0]) / r0[0]:.2e} phase mass drift {(r1[1] - r0[1]) / r0[1]:.2e}")
    assert np.isfinite(got).all()
    for kind, errs in (("max", e), ("rms", rms)):
        for g, names in GROUPS.items():
            worst = max(errs[k] for k in names)
            assert worst <= TOL_SMALL[storage][kind][g], (storage, kind, g, worst)


# 200 steps at 512^3 measured on the MI355X: FP32 mass -1.6e-12, phase mass
# 1.1e-10; BF16 storage mass -3.0e-9, phase mass 2.0e-5 (-1.5e-2 with phi
# rounded to nearest: 8 significant bits, and phi ~ 1 - 1e-3 in the bulk
# rounds the same way at every node); FX16 mass 1.8e-9, phase mass -5.0e-7
@pytest.mark.parametrize("storage,mass_tol,phase_tol", [
    ("fp32", 1e-9, 1e-8),
    ("bf16", 1e-8, 1e-4),
    ("fx16", 1e-8, 1e-5),
])
def test_emulsion_512_long_run(storage, mass_tol, phase_tol):
    case = configs.emulsion()  # 512^3, 200 droplets, radii U[8, 16]
    n = case.shape[0]
    print(f"emulsion 512^3 {storage}: {case.notes}", flush=True)
    with Lattice(n, n, n, storage=storage, **case.lattice) as lat:
        case.apply(lat)
        r0 = lat.reduce()
        lat.step(200)
        r1 = lat.reduce()
    dm = (r1[0] - r0[0]) / r0[0]
    dp = (r1[1] - r0[1]) / r0[1]
    print(f"emulsion 512^3 {storage} 200 steps: mass drift {dm:.2e}, phase mass drift {dp:.2e}, "
          f"KE {r0[2]:.4e} -> {r1[2]:.4e}, enstrophy {r1[6]:.4e}")
    assert np.all(np.isfinite(r1))
    assert abs(dm) <= mass_tol
    assert abs(dp) <= phase_tol
    assert r1[2] > 0
