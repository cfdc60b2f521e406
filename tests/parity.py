"""Parity metric shared by the tests.

Per-field max relative error, each field normalised by the scale of its group:
  rho                      max|d rho| / max|rho|
  u, v, w                  max|d u_a| / max_b max|u_b|      (velocity scale)
  pxx, pyy, pzz, pxy, ...  max|d P_ab| / max_cd max|P_cd|   (stress scale)
  phi                      max|d phi| / max|phi|
A component that is zero in the reference (w of a 2D-like flow) is then judged
against the velocity scale of the flow, not against its own round-off.

rho is also judged against its own fluctuation (`drho_error`): the pressure and
the Laplace jump live in rho - 1, which is ~1e-3 of rho, so a bound relative
to max|rho| alone would let ~1 % errors in the fluctuation pass.
"""
import numpy as np

FIELDS = ("rho", "u", "v", "w", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz", "phi")
GROUP = {0: (0,), 1: (1, 2, 3), 2: (1, 2, 3), 3: (1, 2, 3),
         4: (4, 5, 6, 7, 8, 9), 5: (4, 5, 6, 7, 8, 9), 6: (4, 5, 6, 7, 8, 9),
         7: (4, 5, 6, 7, 8, 9), 8: (4, 5, 6, 7, 8, 9), 9: (4, 5, 6, 7, 8, 9), 10: (10,)}


def field_errors(got, ref, mask=None):
    """dict field -> relative error (got, ref: arrays (11, nz, ny, nx))."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    out = {}
    for q, name in enumerate(FIELDS):
        scale = max(float(np.abs(ref[list(GROUP[q])]).max()), 1e-30)
        d = np.abs(got[q] - ref[q])
        if mask is not None:
            d = d[mask]
        out[name] = float(d.max()) / scale if d.size else 0.0
    return out


def worst(errs, fields=None):
    fields = fields or list(errs)
    return max(errs[f] for f in fields)


def field_rms_errors(got, ref):
    """dict field -> RMS of the difference over the RMS of the field's group."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    out = {}
    for q, name in enumerate(FIELDS):
        scale = max(float(np.sqrt(np.mean(ref[list(GROUP[q])] ** 2))), 1e-30)
        out[name] = float(np.sqrt(np.mean((got[q] - ref[q]) ** 2))) / scale
    return out


def drho_error(got, ref):
    """max|d rho| / max|rho_ref - 1| (the density fluctuation's own scale)."""
    g = np.asarray(got[0], dtype=np.float64)
    r = np.asarray(ref[0], dtype=np.float64)
    return float(np.abs(g - r).max()) / max(float(np.abs(r - 1.0).max()), 1e-30)
