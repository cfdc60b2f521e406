"""The Fortran driver (entropicthreadsafe_amd/fortran/tslb_driver) over the C ABI:
it runs the reference's call sequence and its slice output matches the Python
binding of the same library on the same seeded droplet."""
import os
import subprocess

import numpy as np
import pytest

from entropicthreadsafe_amd import configs
from entropicthreadsafe_amd.tslb import Lattice

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "entropicthreadsafe_amd", "fortran", "tslb_driver")


def test_driver_links_the_library():
    out = subprocess.run(["ldd", DRIVER], capture_output=True, text=True, check=True).stdout
    assert "libtslb.so" in out


@pytest.mark.gpu
def test_driver_droplet_matches_python_binding(tmp_path):
    n = 32
    (tmp_path / "in.nml").write_text(
        "&tslb_run\n nx=32, ny=32, nz=32, ncomp=2, nsteps=20, stamp2D=10,\n"
        " tau=1.0, sigma=0.02, sharp_c=2.1, config='droplet', radius=8.0, noise=1.0e-3\n/\n")
    res = subprocess.run([DRIVER, "in.nml"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(res.stdout[-2000:], res.stderr[-2000:])
    assert res.returncode == 0
    assert "glups:" in res.stdout
    drift = [float(ln.split()[-1]) for ln in res.stdout.splitlines() if ln.startswith("final mass")]
    assert drift and abs(drift[0]) < 1e-6
    # frame 1 is written at step 10 before that step: the state after 9 steps
    xy = np.fromfile(tmp_path / "output" / "out_phixy_000001.raw", dtype=np.float32).reshape(n, n)
    uxz = np.fromfile(tmp_path / "output" / "out_uxz_000001.raw", dtype=np.float32).reshape(n, n)
    assert (tmp_path / "time.dat").exists()
    case = configs.droplet(n=n, radius=8.0, sigma=0.02, beta=0.7, noise=1e-3)
    with Lattice(n, n, n, **case.lattice) as lat:
        case.apply(lat)
        lat.step(9)
        m = lat.download()
    assert np.array_equal(xy, m[10][n // 2 - 1])
    assert np.array_equal(uxz, m[1][:, n // 2 - 1, :])
