"""The bandwidth calibration copy (tslb_stream_copy) runs between steps on the
lattice's own buffers: it copies the current state over the other buffer,
which the next step overwrites, so a run with a calibration in the middle
continues bit for bit as one without it.  The ragged channel lattice covers
the copy's tail (the field stride rounded up to 64 elements)."""
import numpy as np
import pytest

from entropicthreadsafe_amd import configs
from entropicthreadsafe_amd.tslb import Lattice

pytestmark = pytest.mark.gpu

CASES = {
    "droplet-2c-24": ((24, 24, 24), lambda: configs.droplet(n=24, radius=6.0, sigma=0.02, beta=0.7, noise=1e-3)),
    "channel-1c-20x18x22": ((20, 18, 22), lambda: configs.channel(nx=20, ny=18, nz=22)),
}


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_calibration_copy_leaves_the_run_unchanged(storage, name):
    (nx, ny, nz), make = CASES[name]
    case = make()
    runs = []
    for calibrate in (False, True):
        with Lattice(nx, ny, nz, storage=storage, **case.lattice) as lat:
            case.apply(lat)
            lat.step(5)
            if calibrate:
                assert lat.stream_copy_gbs() > 0.0
            lat.step(5)
            runs.append(lat.download())
    assert np.array_equal(runs[0], runs[1])
