"""The C-ABI library: it loads, exports every symbol include/tslb.h declares,
and its host-only entry points behave (no GPU needed)."""
import os
import re
import struct

import numpy as np

from entropicthreadsafe_amd import tslb
from entropicthreadsafe_amd.configs import rand_noseeded_grid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tslb.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tslb_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = tslb.load_library()
    syms = declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), s
        assert s in tslb.SIGNATURES, s


def test_version_and_last_error():
    assert tslb.load_library().tslb_version().startswith(b"tslb")
    assert isinstance(tslb.last_error(), str)


def test_params_layout_matches_header():
    # 10 int32 + 9 float32 = 76 bytes, no padding
    assert tslb.ctypes.sizeof(tslb.Params) == 76
    p = tslb.Params()
    tslb.load_library().tslb_default_params(tslb.ctypes.byref(p))
    assert (p.nx, p.ny, p.nz, p.ncomp) == (16, 16, 16, 1)
    assert abs(p.tau - 0.8) < 1e-7


def test_host_rng_matches_reference_table():
    path = os.path.join(ROOT, "tests", "golden", "rng_golden.txt")
    n = 0
    for ln in open(path):
        if ln.startswith("#"):
            continue
        i, j, k, l, rb, gb = ln.split()
        r = struct.unpack(">f", bytes.fromhex(rb))[0]
        g = struct.unpack(">f", bytes.fromhex(gb))[0]
        assert tslb.rand_noseeded(int(i), int(j), int(k), int(l)) == r
        assert tslb.gauss_noseeded(int(i), int(j), int(k), int(l)) == g
        n += 1
    assert n > 4000


def test_vectorised_rng_matches_library():
    g = rand_noseeded_grid(7, 5, 4, 524)
    for (k, j, i) in [(0, 0, 0), (3, 4, 6), (2, 1, 3), (1, 3, 5)]:
        assert g[k, j, i] == tslb.rand_noseeded(i + 1, j + 1, k + 1, 524)


def test_host_memory_query():
    a = tslb.ctypes.c_double()
    b = tslb.ctypes.c_double()
    assert tslb.load_library().tslb_host_mem_mb(tslb.ctypes.byref(a), tslb.ctypes.byref(b)) == 0
    assert a.value > 0 and b.value > a.value
