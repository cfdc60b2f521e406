"""ctypes wrapper of the CPU oracle (oracle/liboracle_tslb.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the CPU-baseline leg of bench.py, always as the checker.  The product path never
imports this module.

Field order used everywhere (Fortran (1:nx,1:ny,1:nz) arrays, i fastest):
rho, u, v, w, pxx, pyy, pzz, pxy, pxz, pyz, phi.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_tslb.so")

FIELDS = ("rho", "u", "v", "w", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz", "phi")

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle_tslb.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        for sfx, ct in (("f", ctypes.c_float), ("d", ctypes.c_double)):
            cp = ctypes.POINTER(ct)
            getattr(L, f"orc_create_{sfx}").restype = P
            getattr(L, f"orc_create_{sfx}").argtypes = [
                ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                ctypes.POINTER(ctypes.c_int), ctypes.c_double, ctypes.c_double,
                ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                ctypes.c_double]
            getattr(L, f"orc_destroy_{sfx}").argtypes = [P]
            getattr(L, f"orc_set_isfluid_{sfx}").argtypes = [P, ctypes.POINTER(ctypes.c_int32)]
            for nm in ("get_moments", "set_moments", "get_normals"):
                getattr(L, f"orc_{nm}_{sfx}").argtypes = [P, cp]
            getattr(L, f"orc_init_{sfx}").argtypes = [P, cp, cp, cp, cp, cp, ctypes.c_int]
            getattr(L, f"orc_step_{sfx}").argtypes = [P, ctypes.c_int]
            getattr(L, f"orc_pull_step_{sfx}").argtypes = [P, ctypes.c_int]
            getattr(L, f"orc_reduce_{sfx}").argtypes = [P, ctypes.POINTER(ctypes.c_double)]
            getattr(L, f"orc_feq_{sfx}").argtypes = [ct, ct, ct, ct, cp]
            getattr(L, f"orc_aneq_{sfx}").argtypes = [cp, ct, ct, ct, cp]
            getattr(L, f"orc_set_jet_{sfx}").argtypes = [P, ctypes.c_int]
            getattr(L, f"orc_set_zplane_bc_{sfx}").argtypes = [P, cp, cp]
            getattr(L, f"orc_curvature_{sfx}").argtypes = [P, cp]
            getattr(L, f"orc_set_closure_{sfx}").argtypes = [P, cp]
            getattr(L, f"orc_set_colour_{sfx}").argtypes = [P, ctypes.c_double, ctypes.c_double]
            getattr(L, f"orc_colour_values_{sfx}").argtypes = [
                ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, ct, cp, cp, ct, cp,
                cp, cp]
            getattr(L, f"orc_jet_values_{sfx}").argtypes = [ct, ct, cp, cp, cp, cp]
        L.orc_rand_noseeded.restype = ctypes.c_float
        L.orc_rand_noseeded.argtypes = [ctypes.c_int] * 4
        L.orc_gauss_noseeded.restype = ctypes.c_float
        L.orc_gauss_noseeded.argtypes = [ctypes.c_int] * 4
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def rand_noseeded(i, j, k, l):
    return lib().orc_rand_noseeded(int(i), int(j), int(k), int(l))


def gauss_noseeded(i, j, k, l):
    return lib().orc_gauss_noseeded(int(i), int(j), int(k), int(l))


def set_threads(n):
    lib().orc_set_threads(int(n))


def feq(rho, u, v, w, dtype=np.float64):
    sfx, ct = ("d", ctypes.c_double) if dtype == np.float64 else ("f", ctypes.c_float)
    out = np.zeros(27, dtype=dtype)
    getattr(lib(), f"orc_feq_{sfx}")(rho, u, v, w, _ptr(out, ct))
    return out


def aneq(P6, u, v, w, dtype=np.float64):
    sfx, ct = ("d", ctypes.c_double) if dtype == np.float64 else ("f", ctypes.c_float)
    P = np.ascontiguousarray(P6, dtype=dtype)
    out = np.zeros(27, dtype=dtype)
    getattr(lib(), f"orc_aneq_{sfx}")(_ptr(P, ct), u, v, w, _ptr(out, ct))
    return out


def jet_values(omega, rho_y, P_y, u_y, u_x, dtype=np.float64):
    """27 regularized jet-boundary values feq(rho_y, u_x) + (1-omega) w a^neq(P_y, u_y)."""
    sfx, ct = ("d", ctypes.c_double) if dtype == np.float64 else ("f", ctypes.c_float)
    P = np.ascontiguousarray(P_y, dtype=dtype)
    uy = np.ascontiguousarray(u_y, dtype=dtype)
    ux = np.ascontiguousarray(u_x, dtype=dtype)
    out = np.zeros(27, dtype=dtype)
    getattr(lib(), f"orc_jet_values_{sfx}")(omega, rho_y, _ptr(P, ct), _ptr(uy, ct), _ptr(ux, ct),
                                            _ptr(out, ct))
    return out


def colour_values(omega_a, omega_b, sigma, beta, rho, u3, P6, psi, g3, dtype=np.float64):
    """Total and component-A post-collision populations of the colour-gradient
    model for one node (orc_colour_values)."""
    sfx, ct = ("d", ctypes.c_double) if dtype == np.float64 else ("f", ctypes.c_float)
    u = np.ascontiguousarray(u3, dtype=dtype)
    P = np.ascontiguousarray(P6, dtype=dtype)
    g = np.ascontiguousarray(g3, dtype=dtype)
    fo = np.zeros(27, dtype=dtype)
    fa = np.zeros(27, dtype=dtype)
    getattr(lib(), f"orc_colour_values_{sfx}")(omega_a, omega_b, sigma, beta, rho, _ptr(u, ct), _ptr(P, ct),
                                               psi, _ptr(g, ct), _ptr(fo, ct), _ptr(fa, ct))
    return fo, fa


class Oracle:
    """One oracle lattice (population-push form, plus the moment-space pull form).

    Arrays handed in or out are numpy arrays of shape (nz, ny, nx) in C order,
    which is the Fortran (nx, ny, nz) layout with i fastest.
    """

    def __init__(self, nx, ny, nz, ncomp=1, tau=0.6, sigma=0.0, sharp_c=0.0,
                 force=(0.0, 0.0, 0.0), force_mode=0, lin_a=0.0,
                 periodic=(1, 1, 1), dtype=np.float64):
        self.shape = (nz, ny, nx)
        self.n = nx * ny * nz
        self.dtype = np.dtype(dtype)
        self.sfx, self.ct = ("d", ctypes.c_double) if self.dtype == np.float64 else ("f", ctypes.c_float)
        per = (ctypes.c_int * 3)(*[int(p) for p in periodic])
        frc = (ctypes.c_double * 3)(*[float(x) for x in force])
        self._h = getattr(lib(), f"orc_create_{self.sfx}")(
            nx, ny, nz, ncomp, per, float(tau), float(sigma), float(sharp_c), frc,
            int(force_mode), float(lin_a))
        if not self._h:
            raise MemoryError("oracle allocation failed")

    def _fn(self, name):
        return getattr(lib(), f"orc_{name}_{self.sfx}")

    def close(self):
        if self._h:
            self._fn("destroy")(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_isfluid(self, isf):
        a = np.ascontiguousarray(isf, dtype=np.int32).reshape(-1)
        assert a.size == self.n
        self._fn("set_isfluid")(self._h, _ptr(a, ctypes.c_int32))

    def set_colour(self, tau_b, beta):
        """Colour-gradient model (ncomp=3): tau of component B and recolouring beta."""
        self._fn("set_colour")(self._h, float(tau_b), float(beta))

    def set_jet(self, enable=True):
        """Turbulent-jet inlet (k=0) / outlet (k=nz-1) wall planes; set before init."""
        self._fn("set_jet")(self._h, int(bool(enable)))

    def set_zplane_bc(self, w_in, phi_in=None):
        """Moment-space z-plane BC (bcs_TSLB_only_z_turbojet): inflow w (ny, nx) of
        plane k=0, optional phi (two components); applied now and after every
        step.  None disables it."""
        if w_in is None:
            self._fn("set_zplane_bc")(self._h, None, None)
            return
        w = np.ascontiguousarray(w_in, dtype=self.dtype).reshape(-1)
        assert w.size == self.shape[1] * self.shape[2]
        ph = None
        if phi_in is not None:
            ph = np.ascontiguousarray(phi_in, dtype=self.dtype).reshape(-1)
            assert ph.size == w.size
        self._fn("set_zplane_bc")(self._h, _ptr(w, self.ct), None if ph is None else _ptr(ph, self.ct))

    def set_closure(self, clo):
        """Per-node learned-closure fields (7, nz, ny, nx): nu_t, dsigma; None clears."""
        if clo is None:
            self._fn("set_closure")(self._h, None)
            return
        a = np.ascontiguousarray(clo, dtype=self.dtype).reshape(-1)
        assert a.size == 7 * self.n
        self._fn("set_closure")(self._h, _ptr(a, self.ct))

    def init(self, rho, u, v, w, phi=None, prepush=True):
        arrs = [np.ascontiguousarray(x, dtype=self.dtype).reshape(-1) for x in (rho, u, v, w)]
        ph = np.ascontiguousarray(phi if phi is not None else np.zeros(self.shape), dtype=self.dtype).reshape(-1)
        self._fn("init")(self._h, *[_ptr(a, self.ct) for a in arrs], _ptr(ph, self.ct), int(prepush))

    def step(self, n=1):
        self._fn("step")(self._h, int(n))

    def pull_step(self, n=1):
        self._fn("pull_step")(self._h, int(n))

    def moments(self):
        out = np.zeros((11,) + self.shape, dtype=self.dtype)
        self._fn("get_moments")(self._h, _ptr(out, self.ct))
        return out

    def set_moments(self, m):
        a = np.ascontiguousarray(m, dtype=self.dtype)
        assert a.size == 11 * self.n
        self._fn("set_moments")(self._h, _ptr(a, self.ct))

    def normals(self):
        out = np.zeros((3,) + self.shape, dtype=self.dtype)
        self._fn("get_normals")(self._h, _ptr(out, self.ct))
        return out

    def curvature(self):
        """Normals of the current phi, then the gated curvature kappa = -div(n)
        of every node (0 where phi(1-phi)|n|^2 <= 0.2); also refreshes normals()."""
        out = np.zeros(self.shape, dtype=self.dtype)
        self._fn("curvature")(self._h, _ptr(out, self.ct))
        return out

    def reduce(self):
        out = (ctypes.c_double * 7)()
        self._fn("reduce")(self._h, out)
        return np.array(out[:], dtype=np.float64)
