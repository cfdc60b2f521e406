"""CPU oracle of the learned subgrid closure.  TEST INFRASTRUCTURE ONLY: the
product path is the HIP library; only tests and the benchmark's CPU leg use
this module, as the checker.

The reference has no learned closure; the model is defined by the target
description of the TSLB path (a per-node MLP mapping local strain, colour
gradient and phi to an eddy viscosity and six interfacial-stress corrections
of the relaxation, trained on box-filtered fine-grid fields).  Everything
here is restated from that description and from the lattice operators the
reference already has:

* the gradient stencil is the isotropic D3Q27 one of the colour gradient
  (recursive_2c_TSLB3D.f90:555-566), applied to the filtered velocity and
  phi, with the nearest interior node across a wall;
* the box filter is the plain r^3 block mean (Favre-weighted velocity), and
  the subgrid stress is tau_ab = <rho u_a u_b> - rho_bar u~_a u~_b;
* the network is 10 -> 64 -> 64 -> 64 -> 7 with the tanh form of GELU,
  BF16 operands and FP32 accumulation (emulated here: operands are rounded
  to BF16 round-to-nearest-even at the points the GPU rounds them);
* outputs: nu_t = s_0 softplus(o_0), dsigma_k = s_{k+1} o_{k+1};
* the collision uses omega = 1/(tau + 3 nu_t) and adds dsigma to the
  post-collision non-equilibrium stress (the third-order Hermite terms built
  from it as for Pi);
* training loss per sample: sum_k c_k (M_k - T_k)^2 / t_scale^2 with
  M = -2 rho_bar nu_t S + dsigma, T = tau, c = (1,1,1,2,2,2) (Frobenius norm of
  the symmetric tensor), averaged over samples.

Parameter vector layout (float32, natural order, in_features x out_features
row-major per layer): W1 (10x64), b1 (64), W2 (64x64), b2, W3 (64x64), b3,
W4 (64x7), b4 (7), in_scale (10), out_scale (7) -- 9496 floats.
"""
import numpy as np

LAYERS = [(10, 64), (64, 64), (64, 64), (64, 7)]
NFEAT, NOUT = 10, 7
NPARAM = sum(i * o + o for i, o in LAYERS) + NFEAT + NOUT
GELU_C = np.sqrt(2.0 / np.pi)
GELU_A = 0.044715
STRESS_W = np.array([1.0, 1.0, 1.0, 2.0, 2.0, 2.0])


def bf16(x):
    """Round float32 values to BF16 (round to nearest even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(theta):
    """Views of the layers and scales; float64 input stays float64 (finite
    differences), anything else is read as float32 as the GPU holds it."""
    theta = np.asarray(theta)
    if theta.dtype != np.float64:
        theta = theta.astype(np.float32)
    assert theta.size == NPARAM
    out, k = [], 0
    for i, o in LAYERS:
        W = theta[k:k + i * o].reshape(i, o)
        k += i * o
        b = theta[k:k + o]
        k += o
        out.append((W, b))
    in_scale = theta[k:k + NFEAT]
    out_scale = theta[k + NFEAT:k + NFEAT + NOUT]
    return out, in_scale, out_scale


def pack(layers, in_scale, out_scale):
    parts = []
    for W, b in layers:
        parts += [np.asarray(W, np.float32).ravel(), np.asarray(b, np.float32).ravel()]
    parts += [np.asarray(in_scale, np.float32), np.asarray(out_scale, np.float32)]
    return np.concatenate(parts)


def init_params(seed=0, in_scale=None, out_scale=None):
    """He-style random weights, zero biases."""
    rng = np.random.default_rng(seed)
    layers = [(rng.standard_normal((i, o)) * np.sqrt(2.0 / i), np.zeros(o)) for i, o in LAYERS]
    if in_scale is None:
        in_scale = np.ones(NFEAT)
    if out_scale is None:
        out_scale = np.full(NOUT, 1e-3)
    return pack(layers, in_scale, out_scale)


def gelu(z):
    """tanh-form GELU, written as z * sigmoid(2 u), u = c (z + a z^3)."""
    u = GELU_C * (z + GELU_A * z ** 3)
    return z / (1.0 + np.exp(-2.0 * u))


def gelu_grad(z):
    u = GELU_C * (z + GELU_A * z ** 3)
    s = 1.0 / (1.0 + np.exp(-2.0 * u))
    return s + z * s * (1.0 - s) * 2.0 * GELU_C * (1.0 + 3.0 * GELU_A * z ** 2)


def softplus(x):
    return np.logaddexp(0.0, x)


def forward(theta, feats, emulate=True):
    """feats: (n, 10).  Returns (out (n, 7): nu_t, dsigma[6]; cache for backward).
    emulate=False runs the same network in float64 without BF16 rounding."""
    layers, in_scale, out_scale = unpack(theta)
    rnd = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    x = rnd(np.asarray(feats, np.float32) * in_scale[None, :]).astype(np.float64)
    hs, zs = [x], []
    h = x
    for li, (W, b) in enumerate(layers):
        z = h @ rnd(W).astype(np.float64) + b.astype(np.float64)
        if li < 3:
            zs.append(z)
            h = rnd(gelu(z).astype(np.float32) if emulate else gelu(z)).astype(np.float64)
            hs.append(h)
        else:
            o = z
    out = np.empty_like(o)
    out[:, 0] = out_scale[0] * softplus(o[:, 0])
    out[:, 1:] = out_scale[1:] * o[:, 1:]
    return out, (hs, zs, o)


def loss_and_grad(theta, feats, aux, t_scale, emulate=True):
    """Mean loss over samples and its gradient with respect to the weights and
    biases (scales get zero gradient).  aux: (n, 7) = rho_bar, tau[6];
    the strain S is feats[:, :6].  emulate rounds backward operands (deltas,
    weights, activations) to BF16 where the GPU kernel does."""
    layers, in_scale, out_scale = unpack(theta)
    rnd = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    n = feats.shape[0]
    out, (hs, zs, o) = forward(theta, feats, emulate)
    S = np.asarray(feats[:, :6], np.float64)
    rho = np.asarray(aux[:, 0], np.float64)
    T = np.asarray(aux[:, 1:7], np.float64)
    M = -2.0 * rho[:, None] * out[:, :1] * S + out[:, 1:]
    r = M - T
    inv = 1.0 / (t_scale * t_scale * n)
    loss = float(np.sum(STRESS_W[None, :] * r * r) * inv)
    dM = 2.0 * STRESS_W[None, :] * r * inv            # dL/dM
    do = np.empty_like(o)
    dnu = np.sum(dM * (-2.0 * rho[:, None] * S), axis=1)
    do[:, 0] = dnu * out_scale[0] / (1.0 + np.exp(-o[:, 0]))
    do[:, 1:] = dM * out_scale[1:][None, :]
    grads = [None] * 4
    delta = do
    for li in range(3, -1, -1):
        W, b = layers[li]
        d_op = rnd(delta.astype(np.float32) if emulate else delta).astype(np.float64)
        grads[li] = (hs[li].T @ d_op, d_op.sum(axis=0))
        if li > 0:
            back = d_op @ rnd(W).astype(np.float64).T
            delta = back * gelu_grad(zs[li - 1])
    parts = []
    for gw, gb in grads:
        parts += [gw.ravel(), gb.ravel()]
    parts += [np.zeros(NFEAT), np.zeros(NOUT)]
    return loss, np.concatenate(parts)


def iso_grad(f, periodic):
    """Isotropic D3Q27 gradient of f[z, y, x] along x, y, z (3, nz, ny, nx):
    3 sum_l w_l c_l f(x + c_l), nearest interior node across a wall."""
    w = {1: 2.0 / 27.0, 2: 1.0 / 54.0, 3: 1.0 / 216.0}
    nz, ny, nx = f.shape
    pad = np.pad(f, 1, mode="edge")
    for ax, per in zip((2, 1, 0), periodic):   # periodic x, y, z
        if per:
            sl_lo = [slice(None)] * 3
            sl_hi = [slice(None)] * 3
            sl_lo[ax] = 0
            sl_hi[ax] = -1
            src_hi = [slice(None)] * 3
            src_lo = [slice(None)] * 3
            src_hi[ax] = -2
            src_lo[ax] = 1
            pad[tuple(sl_lo)] = pad[tuple(src_hi)]
            pad[tuple(sl_hi)] = pad[tuple(src_lo)]
    g = np.zeros((3,) + f.shape)
    for cz in (-1, 0, 1):
        for cy in (-1, 0, 1):
            for cx in (-1, 0, 1):
                nnz = (cx != 0) + (cy != 0) + (cz != 0)
                if nnz == 0:
                    continue
                v = pad[1 + cz:1 + cz + nz, 1 + cy:1 + cy + ny, 1 + cx:1 + cx + nx]
                wl = 3.0 * w[nnz]
                g[0] += wl * cx * v
                g[1] += wl * cy * v
                g[2] += wl * cz * v
    return g


def features(u, v, w, phi, periodic=(1, 1, 1)):
    """Per-node inputs (nz, ny, nx, 10): strain rate Sxx, Syy, Szz, Sxy, Sxz,
    Syz from the isotropic gradient of the velocity, grad(phi), phi."""
    gu, gv, gw = (iso_grad(np.asarray(a, np.float64), periodic) for a in (u, v, w))
    gp = iso_grad(np.asarray(phi, np.float64), periodic)
    S = [gu[0], gv[1], gw[2], 0.5 * (gu[1] + gv[0]), 0.5 * (gu[2] + gw[0]), 0.5 * (gv[2] + gw[1])]
    return np.stack(S + [gp[0], gp[1], gp[2], np.asarray(phi, np.float64)], axis=-1)


def box_filter(rho, u, v, w, phi, r):
    """r^3 block means: rho_bar, Favre velocity, phi_bar and the subgrid
    stress tau = <rho (u - u~)(u - u~)> (= <rho u u> - rho_bar u~ u~)."""
    def blocks(a):
        nz, ny, nx = a.shape
        return np.asarray(a, np.float64).reshape(nz // r, r, ny // r, r, nx // r, r)

    def mean(a):
        return a.mean(axis=(1, 3, 5))

    R = blocks(rho)
    U, V, W = blocks(u), blocks(v), blocks(w)
    rb = mean(R)
    ut, vt, wt = (mean(R * X) / rb for X in (U, V, W))
    def up(a):
        return a[:, None, :, None, :, None]
    du, dv, dw = U - up(ut), V - up(vt), W - up(wt)
    tau = [mean(R * du * du), mean(R * dv * dv), mean(R * dw * dw),
           mean(R * du * dv), mean(R * du * dw), mean(R * dv * dw)]
    return rb, ut, vt, wt, mean(blocks(phi)), np.stack(tau, axis=-1)


def samples(rho, u, v, w, phi, r, periodic=(1, 1, 1)):
    """Features (n, 10) and aux (n, 7) = rho_bar, tau[6] of the r-filtered field,
    nodes in x-fastest order."""
    rb, ut, vt, wt, pb, tau = box_filter(rho, u, v, w, phi, r)
    f = features(ut, vt, wt, pb, periodic)
    aux = np.concatenate([rb[..., None], tau], axis=-1)
    return f.reshape(-1, NFEAT), aux.reshape(-1, 7)


class Adam:
    """Adam on the weights and biases (the scales are fixed)."""

    def __init__(self, n, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.m = np.zeros(n)
        self.v = np.zeros(n)
        self.t = 0
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps

    def step(self, theta, g):
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        upd = self.lr * mh / (np.sqrt(vh) + self.eps)
        upd[NPARAM - NFEAT - NOUT:] = 0.0
        return (np.asarray(theta, np.float64) - upd).astype(np.float32)
