"""The separated isotropic gradient stencil of the phase-field normals and the
colour gradient (plane_grad / grad_run_step in csrc/tslb.hip), on the CPU.

The reference's K2 stencil (recursive_2c_TSLB3D.f90:555-571) is
    grad phi = 3 (p1 (face+ - face-) + p2 (edges+ - edges-) + p3 (corners+ - corners-))
with p1, p2, p3 = 2/27, 1/54, 1/216, i.e. 3 sum_l w_l c_l phi(x + c_l).  The
kernels use W(dx) W(dy) W(dz) / 216 with W = (1, 4, 1) and form differences of
nearby phi before any weighted sum:
    Px = sum_dy W(dy) (phi(x+1) - phi(x-1)),  Ps = sum W W (phi - phi_centre),
    g_x = (Px(z-1) + 4 Px(z) + Px(z+1)) / 72,
    g_z = (Ps(z+1) - Ps(z-1) + 36 (phi_c(z+1) - phi_c(z-1))) / 72.
Checks: the identity in FP64 (to round-off), and that an FP32 emulation of the
kernels' operation order keeps the unit normal within 1e-6 of FP64 on a
droplet, over an order of magnitude closer than the reference's own
summation order in FP32 (whose error in a bulk phase grows to ~1e-4 once
the field carries noise; tests/test_gpu_parity.py).
"""
import numpy as np

W1 = {1: 2.0 / 27.0, 2: 1.0 / 54.0, 3: 1.0 / 216.0}


def shift(a, dz, dy, dx):
    """a at (z+dz, y+dy, x+dx), periodic."""
    return np.roll(a, (-dz, -dy, -dx), axis=(0, 1, 2))


def ref_gradient(phi, dtype):
    """The reference grouping: per axis, sum each side (faces, edges, corners)
    first, then 3 (p1 (f+ - f-) + p2 (e+ - e-) + p3 (c+ - c-)), in dtype."""
    phi = phi.astype(dtype)
    out = []
    for ax in range(3):
        side = {(s, k): np.zeros_like(phi) for s in (1, -1) for k in (1, 2, 3)}
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    c = (dz, dy, dx)[2 - ax]  # ax 0 = x
                    if c == 0:
                        continue
                    nnz = (dz != 0) + (dy != 0) + (dx != 0)
                    side[(c, nnz)] = (side[(c, nnz)] + shift(phi, dz, dy, dx)).astype(dtype)
        acc = (dtype(W1[1]) * (side[(1, 1)] - side[(-1, 1)])
               + dtype(W1[2]) * (side[(1, 2)] - side[(-1, 2)])
               + dtype(W1[3]) * (side[(1, 3)] - side[(-1, 3)]))
        out.append((dtype(3) * acc).astype(dtype))
    return out  # gx, gy, gz


def separated_gradient(phi, dtype=np.float32):
    """The kernels' operation order (72 grad phi), emulated in dtype with
    fused multiply-adds rounded once."""
    p = phi.astype(dtype)

    def fma_r(a, b, c):
        return (np.asarray(a, np.float64) * np.float64(b) + np.asarray(c, np.float64)).astype(dtype)

    def plane(P):
        a, b, d = shift(P, 0, -1, -1), shift(P, 0, -1, 0), shift(P, 0, -1, 1)
        e, f, g = shift(P, 0, 0, -1), P, shift(P, 0, 0, 1)
        h, i, j = shift(P, 0, 1, -1), shift(P, 0, 1, 0), shift(P, 0, 1, 1)
        x = (d - a) + fma_r(g - e, 4.0, j - h)
        y = (h - a) + fma_r(i - b, 4.0, j - d)
        cm = fma_r(e - f, 4.0, (a - f) + (h - f))
        c0 = (b - f) + (i - f)
        cp = fma_r(g - f, 4.0, (d - f) + (j - f))
        return x, y, fma_r(c0, 4.0, cm + cp), f

    pm, p0, pp = plane(shift(p, -1, 0, 0)), plane(p), plane(shift(p, 1, 0, 0))
    gx = fma_r(p0[0], 4.0, pm[0]) + pp[0]
    gy = fma_r(p0[1], 4.0, pm[1]) + pp[1]
    gz = fma_r(pp[3] - pm[3], 36.0, pp[2] - pm[2])
    return gx, gy, gz


def unit(g, floor):
    gx, gy, gz = [np.asarray(c, np.float64) for c in g]
    inv = 1.0 / (np.sqrt(gx * gx + gy * gy + gz * gz) + floor)
    return np.stack([gx * inv, gy * inv, gz * inv])


def droplet(n=24, radius=6.0, width=2.0):
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    r = np.sqrt((i - n / 2 + 0.3) ** 2 + (j - n / 2 - 0.2) ** 2 + (k - n / 2 + 0.1) ** 2)
    return (0.5 * (1.0 + np.tanh((radius - r) / width))).astype(np.float32)


def test_separated_stencil_is_the_reference_stencil():
    rng = np.random.default_rng(5)
    phi = rng.random((12, 10, 14))
    ref = ref_gradient(phi, np.float64)
    sep = separated_gradient(phi, np.float64)
    for r, s in zip(ref, sep):
        assert np.abs(np.asarray(s, np.float64) / 72.0 - r).max() < 1e-14
    # the weights factor: p1, p2, p3 = W W W / 216 for faces, edges, corners
    assert abs(16.0 / 216.0 - W1[1]) < 1e-17
    assert abs(4.0 / 216.0 - W1[2]) < 1e-17
    assert abs(1.0 / 216.0 - W1[3]) < 1e-17


def test_difference_first_order_is_closer_to_fp64_than_the_reference_order():
    phi = droplet()
    exact = unit(ref_gradient(phi, np.float64), 1e-9)
    sep = unit(separated_gradient(phi), 7.2e-8)      # 72 grad phi: the floor scales too
    ref32 = unit(ref_gradient(phi, np.float32), 1e-9)
    e_sep = np.abs(sep - exact).max()
    e_ref = np.abs(ref32 - exact).max()
    print("unit normal vs FP64: difference-first FP32", e_sep, "reference-order FP32", e_ref)
    assert e_sep < 1e-6
    assert e_ref > 20 * e_sep
