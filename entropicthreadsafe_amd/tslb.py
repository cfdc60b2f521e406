"""Host-side binding of libtslb.so (the C ABI declared in include/tslb.h).

`Lattice` mirrors the reference's driver-level interface: the parameter block
and allocations of recursive_2c_TSLB3D.f90:20-59, the initial state of :87-287,
and one pass of the time loop (:344-840: moments, normals, fused collide +
stream, boundary conditions) as `step()`.  The split entry points `normals()`
and `collide_stream()` correspond to the reference kernels K2 and K3+K1.

The GPU library is the only compute path: if libtslb.so is missing or no HIP
device is present, construction raises.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libtslb.so")

FIELDS = ("rho", "u", "v", "w", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz", "phi")
NFIELDS = 11

FORCE_ADDITIVE = 0
FORCE_SHIFT = 1
STORAGE = {"fp32": 0, "bf16": 1, "fp16": 2, "fx16": 3}


class TslbError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("ncomp", ctypes.c_int32),
        ("periodic", ctypes.c_int32 * 3),
        ("force_mode", ctypes.c_int32),
        ("storage", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("tau", ctypes.c_float),
        ("sigma", ctypes.c_float),
        ("sharp_c", ctypes.c_float),
        ("force", ctypes.c_float * 3),
        ("lin_force", ctypes.c_float),
        ("tau_b", ctypes.c_float),
        ("beta", ctypes.c_float),
    ]


# every symbol include/tslb.h declares, with its ctypes signature
_P = ctypes.c_void_p
_FP = ctypes.POINTER(ctypes.c_float)
_DP = ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "tslb_default_params": (None, [ctypes.POINTER(Params)]),
    "tslb_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.POINTER(Params)]),
    "tslb_destroy": (ctypes.c_int, [_P]),
    "tslb_set_isfluid": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32)]),
    "tslb_upload_moments": (ctypes.c_int, [_P, _FP]),
    "tslb_upload_field": (ctypes.c_int, [_P, ctypes.c_int, _FP]),
    "tslb_download_moments": (ctypes.c_int, [_P, _FP]),
    "tslb_download_field": (ctypes.c_int, [_P, ctypes.c_int, _FP]),
    "tslb_download_normals": (ctypes.c_int, [_P, _FP]),
    "tslb_init_equilibrium": (ctypes.c_int, [_P, ctypes.c_int]),
    "tslb_step": (ctypes.c_int, [_P, ctypes.c_int]),
    "tslb_normals": (ctypes.c_int, [_P]),
    "tslb_collide_stream": (ctypes.c_int, [_P]),
    "tslb_set_zplane_bc": (ctypes.c_int, [_P, _FP, _FP]),
    "tslb_set_jet_bc": (ctypes.c_int, [_P, ctypes.c_int]),
    "tslb_checkpoint_write": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int64]),
    "tslb_checkpoint_read": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]),
    "tslb_upload_plane": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _FP]),
    "tslb_seed_noise": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_float, ctypes.c_int]),
    "tslb_set_storage_range": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_float]),
    "tslb_storage_clipped": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "tslb_reduce": (ctypes.c_int, [_P, _DP]),
    "tslb_reduce_noalloc": (ctypes.c_int, [_P, _P]),
    "tslb_sync": (ctypes.c_int, [_P]),
    "tslb_timing_enable": (ctypes.c_int, [_P, ctypes.c_int]),
    "tslb_timing_read": (ctypes.c_int, [_P, ctypes.c_int, _DP, ctypes.POINTER(ctypes.c_int64)]),
    "tslb_mem_info": (ctypes.c_int, [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    "tslb_host_mem_mb": (ctypes.c_int, [_DP, _DP]),
    "tslb_rand_noseeded": (ctypes.c_float, [ctypes.c_int] * 4),
    "tslb_gauss_noseeded": (ctypes.c_float, [ctypes.c_int] * 4),
    "tslb_stream_copy": (ctypes.c_int, [_P, _DP]),
    "tslb_device_bytes": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_size_t)]),
    "tslb_last_error": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    "tslb_version": (ctypes.c_char_p, []),
    "tslb_closure_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int64]),
    "tslb_closure_destroy": (ctypes.c_int, [_P]),
    "tslb_closure_set_params": (ctypes.c_int, [_P, _FP, ctypes.c_int]),
    "tslb_closure_get_params": (ctypes.c_int, [_P, _FP, ctypes.c_int]),
    "tslb_closure_add_samples": (ctypes.c_int, [_P, _P, ctypes.c_int]),
    "tslb_closure_num_samples": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "tslb_closure_clear_samples": (ctypes.c_int, [_P]),
    "tslb_closure_get_samples": (ctypes.c_int, [_P, _FP, _FP]),
    "tslb_closure_set_samples": (ctypes.c_int, [_P, _FP, _FP, ctypes.c_int64]),
    "tslb_closure_forward": (ctypes.c_int, [_P, _FP]),
    "tslb_closure_loss_grad": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32), ctypes.c_int64, ctypes.c_float,
                                              _DP, _FP]),
    "tslb_closure_train": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                          ctypes.c_uint32, _DP]),
    "tslb_set_closure": (ctypes.c_int, [_P, _P]),
    "tslb_set_closure_fields": (ctypes.c_int, [_P, _FP]),
    "tslb_download_closure": (ctypes.c_int, [_P, _FP]),
}

_lib = None


def load_library(path=LIB_PATH):
    """Load libtslb.so and bind every exported symbol; raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise TslbError(f"{path} is missing: build it with `python -m entropicthreadsafe_amd.build`")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error():
    buf = ctypes.create_string_buffer(1024)
    load_library().tslb_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def _check(rc, what):
    if rc != 0:
        raise TslbError(f"{what} failed ({rc}): {last_error()}")


def rand_noseeded(i, j, k, l):
    return load_library().tslb_rand_noseeded(int(i), int(j), int(k), int(l))


def gauss_noseeded(i, j, k, l):
    return load_library().tslb_gauss_noseeded(int(i), int(j), int(k), int(l))


def mem_info():
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    _check(load_library().tslb_mem_info(ctypes.byref(f), ctypes.byref(t)), "tslb_mem_info")
    return f.value, t.value


def _fptr(a):
    return a.ctypes.data_as(_FP)


class Lattice:
    """A TSLB lattice resident on one GPU.

    Host arrays are numpy float32 arrays of shape (nz, ny, nx) in C order, i.e.
    the reference's Fortran (nx, ny, nz) layout with i fastest.
    """

    def __init__(self, nx, ny, nz, ncomp=1, tau=0.8, sigma=0.0, sharp_c=0.0,
                 force=(0.0, 0.0, 0.0), force_mode=FORCE_ADDITIVE, lin_force=0.0, tau_b=0.0, beta=0.7,
                 periodic=(1, 1, 1), storage="fp32", device=0):
        self.lib = load_library()
        p = Params()
        self.lib.tslb_default_params(ctypes.byref(p))
        p.nx, p.ny, p.nz = int(nx), int(ny), int(nz)
        p.ncomp = int(ncomp)
        for a in range(3):
            p.periodic[a] = int(bool(periodic[a]))
            p.force[a] = float(force[a])
        p.force_mode = int(force_mode)
        p.storage = STORAGE[storage] if isinstance(storage, str) else int(storage)
        p.device = int(device)
        p.tau = float(tau)
        p.sigma = float(sigma)
        p.sharp_c = float(sharp_c)
        p.lin_force = float(lin_force)
        p.tau_b = float(tau_b)
        p.beta = float(beta)
        self.params = p
        self.shape = (int(nz), int(ny), int(nx))
        self.n = int(nx) * int(ny) * int(nz)
        self.ncomp = int(ncomp)
        h = ctypes.c_void_p()
        _check(self.lib.tslb_create(ctypes.byref(h), ctypes.byref(p)), "tslb_create")
        self._h = h

    # -- lifetime -------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self.lib.tslb_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- geometry and state ---------------------------------------------------
    def set_isfluid(self, isfluid):
        if isfluid is None:
            _check(self.lib.tslb_set_isfluid(self._h, None), "tslb_set_isfluid")
            return
        a = np.ascontiguousarray(isfluid, dtype=np.int32).reshape(-1)
        if a.size != self.n:
            raise ValueError("isfluid has the wrong size")
        _check(self.lib.tslb_set_isfluid(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
               "tslb_set_isfluid")

    def upload(self, moments):
        """moments: array (11, nz, ny, nx) or dict field -> (nz, ny, nx)."""
        if isinstance(moments, dict):
            for name, arr in moments.items():
                self.upload_field(name, arr)
            return
        m = np.ascontiguousarray(moments, dtype=np.float32)
        if m.size != NFIELDS * self.n:
            raise ValueError("moments has the wrong size")
        _check(self.lib.tslb_upload_moments(self._h, _fptr(m)), "tslb_upload_moments")

    def upload_field(self, name, arr):
        q = FIELDS.index(name) if isinstance(name, str) else int(name)
        a = np.ascontiguousarray(arr, dtype=np.float32)
        if a.size != self.n:
            raise ValueError(f"field {name} has the wrong size")
        _check(self.lib.tslb_upload_field(self._h, q, _fptr(a)), "tslb_upload_field")

    def download(self):
        out = np.empty((NFIELDS,) + self.shape, dtype=np.float32)
        _check(self.lib.tslb_download_moments(self._h, _fptr(out)), "tslb_download_moments")
        return out

    def download_field(self, name):
        q = FIELDS.index(name) if isinstance(name, str) else int(name)
        out = np.empty(self.shape, dtype=np.float32)
        _check(self.lib.tslb_download_field(self._h, q, _fptr(out)), "tslb_download_field")
        return out

    def download_normals(self):
        out = np.empty((3,) + self.shape, dtype=np.float32)
        _check(self.lib.tslb_download_normals(self._h, _fptr(out)), "tslb_download_normals")
        return out

    def init_equilibrium(self, prepush_phi=True):
        _check(self.lib.tslb_init_equilibrium(self._h, int(bool(prepush_phi))), "tslb_init_equilibrium")

    def set_zplane_bc(self, w_in, phi_in=None):
        if w_in is None:
            _check(self.lib.tslb_set_zplane_bc(self._h, None, None), "tslb_set_zplane_bc")
            return
        w = np.ascontiguousarray(w_in, dtype=np.float32).reshape(-1)
        ph = None if phi_in is None else np.ascontiguousarray(phi_in, dtype=np.float32).reshape(-1)
        plane = self.shape[1] * self.shape[2]
        if w.size != plane or (ph is not None and ph.size != plane):
            raise ValueError("inflow planes must hold nx*ny values")
        _check(self.lib.tslb_set_zplane_bc(self._h, _fptr(w), None if ph is None else _fptr(ph)),
               "tslb_set_zplane_bc")

    def checkpoint_write(self, path, step=0):
        """Stored state verbatim (exact restart); see tslb_checkpoint_write."""
        _check(self.lib.tslb_checkpoint_write(self._h, os.fsencode(path), int(step)),
               "tslb_checkpoint_write")

    def checkpoint_read(self, path):
        """Restore a checkpoint written for the same lattice; returns its step."""
        st = ctypes.c_int64(0)
        _check(self.lib.tslb_checkpoint_read(self._h, os.fsencode(path), ctypes.byref(st)),
               "tslb_checkpoint_read")
        return int(st.value)

    def set_jet_bc(self, enable=True):
        """Turbulent-jet inlet (plane 0) / outlet (plane nz-1) walls; see
        tslb_set_jet_bc.  Call after set_isfluid and before init_equilibrium."""
        _check(self.lib.tslb_set_jet_bc(self._h, int(bool(enable))), "tslb_set_jet_bc")

    def upload_plane(self, name, k, arr):
        """One z-plane k (0-based) of a field, arr of shape (ny, nx)."""
        q = FIELDS.index(name) if isinstance(name, str) else int(name)
        a = np.ascontiguousarray(arr, dtype=np.float32).reshape(-1)
        if a.size != self.shape[1] * self.shape[2]:
            raise ValueError("plane has the wrong size")
        _check(self.lib.tslb_upload_plane(self._h, q, int(k), _fptr(a)), "tslb_upload_plane")

    def set_storage_range(self, name, half_range):
        """FX16 storage: a half range for the field's group (rho; u, v, w; the
        six Pi^neq; phi), rounded up to the range 32767 * 2^-m of a
        power-of-two step; values beyond saturate, the group is re-encoded."""
        q = FIELDS.index(name) if isinstance(name, str) else int(name)
        _check(self.lib.tslb_set_storage_range(self._h, q, float(half_range)), "tslb_set_storage_range")

    def storage_clipped(self):
        """FX16 storage: values clipped at their range since creation."""
        n = ctypes.c_int64()
        _check(self.lib.tslb_storage_clipped(self._h, ctypes.byref(n)), "tslb_storage_clipped")
        return n.value

    def seed_noise(self, name, amp, l):
        """field += amp*(2 r - 1), r = rand_noseeded(i, j, k, l), on the device."""
        q = FIELDS.index(name) if isinstance(name, str) else int(name)
        _check(self.lib.tslb_seed_noise(self._h, q, float(amp), int(l)), "tslb_seed_noise")

    # -- stepping ---------------------------------------------------------------
    def step(self, nsteps=1):
        _check(self.lib.tslb_step(self._h, int(nsteps)), "tslb_step")

    def normals(self):
        _check(self.lib.tslb_normals(self._h), "tslb_normals")

    def collide_stream(self):
        _check(self.lib.tslb_collide_stream(self._h), "tslb_collide_stream")

    def sync(self):
        _check(self.lib.tslb_sync(self._h), "tslb_sync")

    # -- diagnostics ------------------------------------------------------------
    def reduce(self):
        out = (ctypes.c_double * 7)()
        _check(self.lib.tslb_reduce(self._h, out), "tslb_reduce")
        return np.array(out[:], dtype=np.float64)

    def timing_enable(self, on=True):
        _check(self.lib.tslb_timing_enable(self._h, int(bool(on))), "tslb_timing_enable")

    def timing_read(self, kernel=0):
        ms, cnt = ctypes.c_double(), ctypes.c_int64()
        _check(self.lib.tslb_timing_read(self._h, int(kernel), ctypes.byref(ms), ctypes.byref(cnt)),
               "tslb_timing_read")
        return ms.value, cnt.value

    def stream_copy_gbs(self):
        """Bandwidth of a plain copy kernel over the moment buffer (GB/s)."""
        g = ctypes.c_double()
        _check(self.lib.tslb_stream_copy(self._h, ctypes.byref(g)), "tslb_stream_copy")
        return g.value

    def device_bytes(self):
        b = ctypes.c_size_t()
        _check(self.lib.tslb_device_bytes(self._h, ctypes.byref(b)), "tslb_device_bytes")
        return b.value

    # -- learned closure ------------------------------------------------------
    def set_closure(self, model):
        """Attach a Closure (recomputed before every collision); None detaches."""
        _check(self.lib.tslb_set_closure(self._h, model._h if model is not None else None), "tslb_set_closure")
        self._closure = model

    def set_closure_fields(self, clo):
        """Fixed per-node closure fields clo (7, nz, ny, nx): nu_t, dsigma; None detaches."""
        if clo is None:
            _check(self.lib.tslb_set_closure_fields(self._h, None), "tslb_set_closure_fields")
            return
        a = np.ascontiguousarray(clo, dtype=np.float32)
        if a.size != 7 * self.n:
            raise ValueError("closure fields have the wrong size")
        _check(self.lib.tslb_set_closure_fields(self._h, _fptr(a)), "tslb_set_closure_fields")

    def download_closure(self):
        out = np.empty((7,) + self.shape, dtype=np.float32)
        _check(self.lib.tslb_download_closure(self._h, _fptr(out)), "tslb_download_closure")
        return out


CLOSURE_NPARAM = 9496


class Closure:
    """The learned subgrid closure (tslb_closure_* of include/tslb.h): an MLP
    10 -> 64 -> 64 -> 64 -> 7 on the matrix cores, its training samples
    (box-filtered lattice states) and its Adam state, on the current device."""

    def __init__(self, max_samples=0):
        self.lib = load_library()
        h = _P()
        _check(self.lib.tslb_closure_create(ctypes.byref(h), int(max_samples)), "tslb_closure_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self.lib.tslb_closure_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, theta):
        a = np.ascontiguousarray(theta, dtype=np.float32)
        if a.size != CLOSURE_NPARAM:
            raise ValueError("expected %d parameters" % CLOSURE_NPARAM)
        _check(self.lib.tslb_closure_set_params(self._h, _fptr(a), CLOSURE_NPARAM), "tslb_closure_set_params")

    def get_params(self):
        out = np.empty(CLOSURE_NPARAM, dtype=np.float32)
        _check(self.lib.tslb_closure_get_params(self._h, _fptr(out), CLOSURE_NPARAM), "tslb_closure_get_params")
        return out

    def add_samples(self, lattice, factor):
        _check(self.lib.tslb_closure_add_samples(self._h, lattice._h, int(factor)), "tslb_closure_add_samples")

    @property
    def num_samples(self):
        n = ctypes.c_int64()
        _check(self.lib.tslb_closure_num_samples(self._h, ctypes.byref(n)), "tslb_closure_num_samples")
        return n.value

    def clear_samples(self):
        _check(self.lib.tslb_closure_clear_samples(self._h), "tslb_closure_clear_samples")

    def get_samples(self):
        """(feat (10, ns), aux (7, ns)) of the stored samples."""
        ns = self.num_samples
        f = np.empty((10, ns), dtype=np.float32)
        a = np.empty((7, ns), dtype=np.float32)
        _check(self.lib.tslb_closure_get_samples(self._h, _fptr(f), _fptr(a)), "tslb_closure_get_samples")
        return f, a

    def set_samples(self, feat, aux):
        f = np.ascontiguousarray(feat, dtype=np.float32)
        a = np.ascontiguousarray(aux, dtype=np.float32)
        ns = f.shape[1]
        if f.shape != (10, ns) or a.shape != (7, ns):
            raise ValueError("feat must be (10, ns) and aux (7, ns)")
        _check(self.lib.tslb_closure_set_samples(self._h, _fptr(f), _fptr(a), ns), "tslb_closure_set_samples")

    def forward(self):
        """Outputs (7, ns) of the stored samples: nu_t, dsigma."""
        out = np.empty((7, self.num_samples), dtype=np.float32)
        _check(self.lib.tslb_closure_forward(self._h, _fptr(out)), "tslb_closure_forward")
        return out

    def loss_grad(self, t_scale, idx=None, grad=True):
        loss = ctypes.c_double()
        g = np.empty(CLOSURE_NPARAM, dtype=np.float32) if grad else None
        if idx is None:
            ip, nb = None, 0
        else:
            ia = np.ascontiguousarray(idx, dtype=np.int32)
            ip, nb = ia.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ia.size
        _check(self.lib.tslb_closure_loss_grad(self._h, ip, nb, float(t_scale), ctypes.byref(loss),
                                               _fptr(g) if grad else None), "tslb_closure_loss_grad")
        return (loss.value, g) if grad else loss.value

    def train(self, steps, batch, lr, t_scale, seed=0):
        loss = ctypes.c_double()
        _check(self.lib.tslb_closure_train(self._h, int(steps), int(batch), float(lr), float(t_scale),
                                           int(seed) & 0xFFFFFFFF, ctypes.byref(loss)), "tslb_closure_train")
        return loss.value
