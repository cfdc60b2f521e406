"""tslb_reduce_noalloc (sums written to device memory; partial sums in the
library's fixed device array) against tslb_reduce on the same state, called
through the C ABI: two calls in a row, the second on a smaller lattice, and
a third call of the first size after it (the array keeps the last call's
partials, which must not leak into the next)."""
import ctypes

import numpy as np
import pytest

from entropicthreadsafe_amd import configs
from entropicthreadsafe_amd.tslb import Lattice, load_library

pytestmark = pytest.mark.gpu


def _noalloc(lat, out):
    rc = load_library().tslb_reduce_noalloc(lat._h, ctypes.c_void_p(out.data_ptr()))
    assert rc == 0
    lat.sync()
    return out.cpu().numpy()


def test_noalloc_reduce_matches_reduce():
    torch = pytest.importorskip("torch")
    out = torch.full((7,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    for n, steps, ncomp in ((72, 6, 2), (20, 3, 1), (72, 2, 2)):
        if ncomp == 2:
            case = configs.droplet(n=n, radius=n / 5, noise=1e-3)
        else:
            case = configs.tgv(n=n)
        with Lattice(n, n, n, **case.lattice) as lat:
            case.apply(lat)
            lat.step(steps)
            got = _noalloc(lat, out)
            ref = lat.reduce()
        print(n, got, ref)
        assert np.all(np.isfinite(got))
        assert np.array_equal(got, ref), (n, got, ref)


def test_noalloc_rejects_a_null_output():
    with Lattice(8, 8, 8) as lat:
        assert load_library().tslb_reduce_noalloc(lat._h, None) != 0
