"""Moment-space checkpoint (tslb_checkpoint_write / _read): the stored state
is saved verbatim, so a restarted run continues bit for bit."""
import numpy as np
import pytest

from entropicthreadsafe_amd import configs
from entropicthreadsafe_amd.tslb import Lattice, TslbError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage", ["fp32", "fp16", "bf16", "fx16"])
def test_restart_is_bitwise(tmp_path, storage):
    case = configs.droplet(n=24, radius=6.0, sigma=0.02, beta=0.7, noise=1e-3)
    ck = str(tmp_path / "state.ckpt")
    with Lattice(24, 24, 24, storage=storage, **case.lattice) as lat:
        if storage == "fx16":
            # a non-default range travels with the checkpoint
            lat.set_storage_range("u", 0.05)
        case.apply(lat)
        lat.step(7)
        lat.checkpoint_write(ck, step=7)
        lat.step(9)
        straight = lat.download()
    with Lattice(24, 24, 24, storage=storage, **case.lattice) as lat:
        assert lat.checkpoint_read(ck) == 7
        lat.step(9)
        restarted = lat.download()
    assert np.array_equal(straight, restarted)


def test_checkpoint_rejects_another_lattice(tmp_path):
    case = configs.tgv(n=16)
    ck = str(tmp_path / "tgv.ckpt")
    with Lattice(16, 16, 16, **case.lattice) as lat:
        case.apply(lat)
        lat.checkpoint_write(ck)
    with Lattice(16, 16, 32, **case.lattice) as lat:
        with pytest.raises(TslbError):
            lat.checkpoint_read(ck)
    with Lattice(16, 16, 16, storage="fp16", **case.lattice) as lat:
        with pytest.raises(TslbError):
            lat.checkpoint_read(ck)
    (tmp_path / "junk.ckpt").write_bytes(b"not a checkpoint" * 8)
    with Lattice(16, 16, 16, **case.lattice) as lat:
        with pytest.raises(TslbError):
            lat.checkpoint_read(str(tmp_path / "junk.ckpt"))


def test_checkpoint_earlier_formats(tmp_path):
    """A version 1 file of FP16 storage (header, then the fields: no scales
    and no step count) is still read, and holds the same state; an FX16
    file of the earlier range-scaled format (version 3) is refused."""
    import struct
    case = configs.droplet(n=16, radius=4.0, sigma=0.02, beta=0.7, noise=1e-3)
    ck = tmp_path / "state.ckpt"
    with Lattice(16, 16, 16, storage="fp16", **case.lattice) as lat:
        case.apply(lat)
        lat.step(3)
        lat.checkpoint_write(str(ck), step=3)
        ref = lat.download()
    raw = ck.read_bytes()
    assert struct.unpack_from("<i", raw, 8)[0] == 4
    # version 1: the 64-byte header, then the fields (drop 4 scales + step)
    v1 = bytearray(raw[:64] + raw[64 + 20:])
    struct.pack_into("<i", v1, 8, 1)
    (tmp_path / "v1.ckpt").write_bytes(bytes(v1))
    with Lattice(16, 16, 16, storage="fp16", **case.lattice) as lat:
        assert lat.checkpoint_read(str(tmp_path / "v1.ckpt")) == 3
        assert np.array_equal(lat.download(), ref)
    with Lattice(16, 16, 16, storage="fx16", **case.lattice) as lat:
        case.apply(lat)
        lat.checkpoint_write(str(ck))
    v3 = bytearray(ck.read_bytes())
    struct.pack_into("<i", v3, 8, 3)
    (tmp_path / "v3.ckpt").write_bytes(bytes(v3))
    with Lattice(16, 16, 16, storage="fx16", **case.lattice) as lat:
        with pytest.raises(TslbError, match="version 3"):
            lat.checkpoint_read(str(tmp_path / "v3.ckpt"))
