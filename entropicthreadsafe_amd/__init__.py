"""MI355X-native TSLB: the thread-safe recursive-regularized D3Q27 lattice
Boltzmann step for one- and two-component flow on one GPU.

The compute path is libtslb.so (hand-written HIP kernels for gfx950) behind
the C ABI of include/tslb.h; `tslb.Lattice` is its host-side binding.
"""
from .tslb import FIELDS, Lattice, TslbError, load_library  # noqa: F401

__all__ = ["FIELDS", "Lattice", "TslbError", "load_library"]
