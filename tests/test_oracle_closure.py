"""Oracle pins for the learned subgrid closure (oracle/closure.py and the
closure fields of the C oracle's collision).

The closure has no reference counterpart; its oracle is pinned against what
the mathematics fixes and against independent implementations: torch's BF16
rounding and tanh-GELU MLP, finite differences of the loss, the exactness of
the isotropic gradient on linear fields, conservation by the box filter, and
the reduction of the closure collision to a plain collision with a shifted
relaxation time.
"""
import numpy as np
import pytest
import torch

import closure as cl
import pyoracle as po


def _params(seed=1):
    return cl.init_params(seed, in_scale=np.full(10, 2.0), out_scale=np.array([1e-2] + [1e-3] * 6))


def _torch_forward(theta, F):
    layers, ins, outs = cl.unpack(theta)
    x = torch.tensor(np.asarray(F, np.float64) * ins.astype(np.float64))
    for i, (W, b) in enumerate(layers):
        x = x @ torch.tensor(W.astype(np.float64)) + torch.tensor(b.astype(np.float64))
        if i < 3:
            x = torch.nn.functional.gelu(x, approximate="tanh")
    o = x.numpy()
    return np.concatenate([outs[0] * np.logaddexp(0, o[:, :1]), outs[1:] * o[:, 1:]], axis=1)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(10000) * 10.0 ** rng.integers(-30, 30, 10000),
                        1 + np.arange(-64, 64) * 2.0 ** -9]).astype(np.float32)
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(cl.bf16(x), ref)


def test_forward_matches_torch_mlp():
    th = _params()
    F = (np.random.default_rng(2).standard_normal((200, 10)) * 0.3).astype(np.float32)
    out, _ = cl.forward(th, F, emulate=False)
    ref = _torch_forward(th, F)
    assert np.abs(out - ref).max() < 1e-14
    # the BF16-operand network stays within BF16 accuracy of it
    oute, _ = cl.forward(th, F, emulate=True)
    assert np.abs(oute - ref).max() < 2e-2 * np.abs(ref).max()


def _batch(n=64, seed=3):
    rng = np.random.default_rng(seed)
    F = (rng.standard_normal((n, 10)) * 0.3).astype(np.float32)
    aux = np.concatenate([1 + 0.01 * rng.standard_normal((n, 1)), 1e-3 * rng.standard_normal((n, 6))], axis=1)
    return F, aux


def test_gradient_matches_finite_differences():
    th = _params()
    F, aux = _batch()
    L, g = cl.loss_and_grad(th, F, aux, 1e-3, emulate=False)
    eps = 1e-4
    for k in (0, 5, 639, 640, 700, 4801, 5000, 8960, 9030, 9472, 9478):
        tp = th.astype(np.float64).copy()
        tm = tp.copy()
        tp[k] += eps
        tm[k] -= eps
        fd = (cl.loss_and_grad(tp, F, aux, 1e-3, emulate=False)[0]
              - cl.loss_and_grad(tm, F, aux, 1e-3, emulate=False)[0]) / (2 * eps)
        assert abs(g[k] - fd) <= 1e-5 * max(1.0, abs(fd)), (k, g[k], fd)
    assert np.all(g[cl.NPARAM - 17:] == 0)


def test_gradient_matches_torch_autograd():
    th = _params(4)
    F, aux = _batch(32, 5)
    L, g = cl.loss_and_grad(th, F, aux, 2e-3, emulate=False)
    layers, ins, outs = cl.unpack(th)
    ps = []
    x = torch.tensor(np.asarray(F, np.float64) * ins.astype(np.float64))
    for i, (W, b) in enumerate(layers):
        Wt = torch.tensor(W.astype(np.float64), requires_grad=True)
        bt = torch.tensor(b.astype(np.float64), requires_grad=True)
        ps += [Wt, bt]
        x = x @ Wt + bt
        if i < 3:
            x = torch.nn.functional.gelu(x, approximate="tanh")
    nu = float(outs[0]) * torch.nn.functional.softplus(x[:, 0])
    ds = x[:, 1:] * torch.tensor(outs[1:].astype(np.float64))
    S = torch.tensor(F[:, :6].astype(np.float64))
    rho = torch.tensor(aux[:, 0])
    M = -2.0 * rho[:, None] * nu[:, None] * S + ds
    r = M - torch.tensor(aux[:, 1:7])
    loss = (torch.tensor(cl.STRESS_W) * r * r).sum() / (2e-3 ** 2 * F.shape[0])
    loss.backward()
    gt = np.concatenate([p.grad.numpy().ravel() for p in ps])
    assert abs(loss.item() - L) < 1e-12 * abs(L)
    assert np.abs(g[:gt.size] - gt).max() < 1e-10 * np.abs(gt).max()


def test_bf16_gradient_close_to_exact():
    th = _params()
    F, aux = _batch(256, 6)
    _, g = cl.loss_and_grad(th, F, aux, 1e-3, emulate=False)
    _, ge = cl.loss_and_grad(th, F, aux, 1e-3, emulate=True)
    assert np.linalg.norm(ge - g) < 2e-2 * np.linalg.norm(g)


def test_features_exact_on_linear_fields():
    """The isotropic gradient is exact for linear fields (away from walls)."""
    n = 12
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    A = np.array([[0.01, -0.02, 0.005], [0.003, 0.0, -0.004], [0.002, 0.006, -0.01]])  # du_a/dx_b
    u = A[0, 0] * i + A[0, 1] * j + A[0, 2] * k
    v = A[1, 0] * i + A[1, 1] * j + A[1, 2] * k
    w = A[2, 0] * i + A[2, 1] * j + A[2, 2] * k
    phi = 0.3 + 0.01 * i - 0.02 * k
    f = cl.features(u, v, w, phi, periodic=(0, 0, 0))
    inner = f[1:-1, 1:-1, 1:-1]
    S = 0.5 * (A + A.T)
    exp = [S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2], 0.01, 0.0, -0.02]
    for q, e in enumerate(exp):
        assert np.abs(inner[..., q] - e).max() < 1e-14, q
    assert np.abs(inner[..., 9] - phi[1:-1, 1:-1, 1:-1]).max() == 0.0
    # across a wall the stencil reads the nearest interior node (one-sided)
    assert abs(f[0, 5, 5, 8] - (-0.02) * 0.5) < 1e-14


def test_features_periodic_wave():
    """A periodic sine: the stencil gives sin(k)/k times the exact derivative
    (its symbol along an axis is 3 * 2 * (w1 + 4 w7 + 4 w19) sin(k) = sin(k))."""
    n = 16
    kk = 2 * np.pi / n
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    u = np.sin(kk * j)
    f = cl.features(u, 0 * u, 0 * u, 0 * u)
    exp = 0.5 * np.sin(kk) * np.cos(kk * j)   # S_xy = du/dy / 2
    assert np.abs(f[..., 3] - exp).max() < 1e-14


def test_box_filter_conserves_and_gives_the_subgrid_stress():
    rng = np.random.default_rng(7)
    n, r = 16, 4
    rho = 1 + 0.01 * rng.standard_normal((n, n, n))
    u, v, w = (0.02 * rng.standard_normal((n, n, n)) for _ in range(3))
    phi = rng.uniform(0, 1, (n, n, n))
    rb, ut, vt, wt, pb, tau = cl.box_filter(rho, u, v, w, phi, r)
    assert abs(rb.sum() * r ** 3 - rho.sum()) < 1e-10
    assert abs((rb * ut).sum() * r ** 3 - (rho * u).sum()) < 1e-12
    assert abs(pb.sum() * r ** 3 - phi.sum()) < 1e-10
    # tau = <rho u u> - rho_bar u~ u~ (the definition, block by block)
    blk = (slice(4, 8), slice(8, 12), slice(0, 4))
    R, U, V = rho[blk], u[blk], v[blk]
    exp_xy = (R * U * V).mean() - R.mean() * ((R * U).mean() / R.mean()) * ((R * V).mean() / R.mean())
    assert abs(tau[1, 2, 0, 3] - exp_xy) < 1e-15
    # positive semidefinite
    for a in range(3):
        assert (tau[..., a] >= 0).all()
    # a uniform velocity has no subgrid stress
    _, _, _, _, _, t0 = cl.box_filter(rho, 0 * u + 0.03, 0 * v - 0.01, 0 * w, phi, r)
    assert np.abs(t0).max() < 1e-17


def _fields(n, seed):
    rng = np.random.default_rng(seed)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    rho = 1 + 1e-3 * rng.standard_normal((n, n, n))
    u = 0.03 * np.sin(2 * np.pi * j / n) + 1e-3 * rng.standard_normal((n, n, n))
    v = 0.02 * np.cos(2 * np.pi * k / n)
    w = 0.01 * np.sin(2 * np.pi * i / n)
    return rho, u, v, w


def test_closure_collision_reduces_to_shifted_tau():
    """Uniform nu_t and no stress correction: the collision is the plain one
    with tau' = tau + 3 nu_t; zero fields reproduce the plain step (to
    rounding: (1 - omega) multiplies Pi before the Hermite terms)."""
    n, tau, nut = 12, 0.6, 0.02
    fields = _fields(n, 1)
    runs = []
    for kind in ("zero", "plain", "nut", "shifted"):
        o = po.Oracle(n, n, n, ncomp=1, tau=tau + 3 * nut if kind == "shifted" else tau)
        if kind in ("zero", "nut"):
            clo = np.zeros((7, n, n, n))
            if kind == "nut":
                clo[0] = nut
            o.set_closure(clo)
        o.init(*fields)
        o.pull_step(6)
        runs.append(o.moments())
        o.close()
    assert np.abs(runs[0] - runs[1]).max() < 1e-15
    assert np.abs(runs[2] - runs[3]).max() < 1e-15


def test_closure_stress_keeps_mass_and_momentum():
    """dsigma enters the second moment only: total mass and momentum are
    unchanged by it on a periodic box (no force)."""
    n = 12
    fields = _fields(n, 2)
    rng = np.random.default_rng(3)
    clo = np.zeros((7, n, n, n))
    clo[0] = 0.01 * rng.uniform(size=(n, n, n))
    clo[1:] = 1e-4 * rng.standard_normal((6, n, n, n))
    o = po.Oracle(n, n, n, ncomp=1, tau=0.7)
    o.set_closure(clo)
    o.init(*fields)
    r0 = o.moments()
    o.step(5)
    r1 = o.moments()
    o.close()
    m0 = r0[0].sum()
    assert abs(r1[0].sum() - m0) < 1e-11 * m0
    for a in (1, 2, 3):
        assert abs((r1[0] * r1[a]).sum() - (r0[0] * r0[a]).sum()) < 1e-12
    # and it changes the state
    o = po.Oracle(n, n, n, ncomp=1, tau=0.7)
    o.init(*fields)
    o.step(5)
    assert np.abs(o.moments() - r1).max() > 1e-6
    o.close()


@pytest.mark.parametrize("form", ["push", "pull"])
def test_closure_push_and_pull_agree(form):
    n = 10
    fields = _fields(n, 4)
    rng = np.random.default_rng(5)
    clo = np.zeros((7, n, n, n))
    clo[0] = 0.01 * rng.uniform(size=(n, n, n))
    clo[1:] = 1e-4 * rng.standard_normal((6, n, n, n))
    res = []
    for f in ("push", "pull"):
        o = po.Oracle(n, n, n, ncomp=1, tau=0.7, periodic=(1, 0, 1))
        isf = np.ones((n, n, n), np.int32)
        isf[3:5, 4:6, 2:7] = 0
        o.set_isfluid(isf)
        o.set_closure(clo)
        o.init(*fields)
        (o.step if f == "push" else o.pull_step)(4)
        res.append(o.moments())
        o.close()
    assert np.array_equal(res[0], res[1])
