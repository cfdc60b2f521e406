"""The learned closure on the GPU (tslb_closure_*, k_mlp_*, k_features,
k_box_filter, the closure variant of k_fused) against oracle/closure.py and
the C oracle's closure collision.

Tolerances: the MLP runs on BF16 matrix cores with FP32 accumulation, the
oracle emulates the same BF16 roundings with FP64 accumulation, so outputs
agree to within a few BF16 roundings of single activations (rel. 2e-3 of the
output scale); gradients to 2 %.  The collision with given closure fields
is FP32 against FP64: per-field max relative error <= 1e-5 as every parity
test of the step.
"""
import numpy as np
import pytest

import closure as cl
import pyoracle as po
from entropicthreadsafe_amd.tslb import Closure, Lattice, TslbError
from parity import field_errors

pytestmark = pytest.mark.gpu


def _params(seed=1, out0=1e-2):
    return cl.init_params(seed, in_scale=np.full(10, 2.0), out_scale=np.array([out0] + [1e-3] * 6))


def _samples(ns, seed=3):
    rng = np.random.default_rng(seed)
    F = (rng.standard_normal((ns, 10)) * 0.3).astype(np.float32)
    aux = np.concatenate([1 + 0.01 * rng.standard_normal((ns, 1)), 1e-3 * rng.standard_normal((ns, 6))],
                         axis=1).astype(np.float32)
    return F, aux


@pytest.mark.parametrize("ns", [1, 16, 1000, 40000])
def test_mlp_forward_matches_oracle(ns):
    th = _params()
    F, aux = _samples(ns)
    with Closure(max_samples=ns) as m:
        m.set_params(th)
        assert np.array_equal(m.get_params(), th)
        m.set_samples(F.T, aux.T)
        got = m.forward().T
    ref, _ = cl.forward(th, F, emulate=True)
    scale = np.abs(ref).max(axis=0)
    err = np.abs(got - ref).max(axis=0) / scale
    print("forward rel err per output", err)
    assert np.isfinite(got).all()
    assert (err < 2e-3).all()
    assert np.all(got[:, 0] >= 0)


@pytest.mark.parametrize("nb", [16, 333, 5000])
def test_mlp_loss_and_gradient_match_oracle(nb):
    th = _params(2)
    F, aux = _samples(6000, 4)
    idx = np.random.default_rng(5).integers(0, 6000, nb).astype(np.int32)
    with Closure(max_samples=6000) as m:
        m.set_params(th)
        m.set_samples(F.T, aux.T)
        loss, g = m.loss_grad(1e-3, idx=idx)
    lref, gref = cl.loss_and_grad(th, F[idx], aux[idx].astype(np.float64), 1e-3, emulate=True)
    print("loss", loss, lref, "grad rel", np.linalg.norm(g - gref) / np.linalg.norm(gref))
    assert abs(loss - lref) < 1e-3 * lref
    assert np.linalg.norm(g - gref) < 2e-2 * np.linalg.norm(gref)
    # per layer too
    for a, b in ((0, 640), (640, 704), (704, 4800), (4864, 8960), (9024, 9472), (9472, 9479)):
        assert np.linalg.norm(g[a:b] - gref[a:b]) < 5e-2 * np.linalg.norm(gref[a:b]) + 1e-30, (a, b)
    assert np.all(g[9479:] == 0)


def test_training_reduces_loss_and_finds_the_eddy_viscosity():
    """Targets from an eddy-viscosity law tau = -2 rho nu* S: Adam on the GPU
    drives the loss down and nu_t towards nu*."""
    rng = np.random.default_rng(6)
    ns, nu_star = 20000, 4e-3
    F = np.zeros((ns, 10), np.float32)
    F[:, :6] = 0.02 * rng.standard_normal((ns, 6))
    F[:, 6:9] = 0.1 * rng.standard_normal((ns, 3))
    F[:, 9] = rng.uniform(0, 1, ns)
    rho = 1 + 0.01 * rng.standard_normal(ns)
    aux = np.concatenate([rho[:, None], -2 * rho[:, None] * nu_star * F[:, :6]], axis=1).astype(np.float32)
    t_scale = float(np.sqrt((aux[:, 1:] ** 2).mean()))
    th = cl.init_params(7, in_scale=np.array([50.0] * 6 + [10.0] * 3 + [2.0]),
                        out_scale=np.array([1e-2] + [1e-4] * 6))
    with Closure(max_samples=ns) as m:
        m.set_params(th)
        m.set_samples(F.T, aux.T)
        l0 = m.loss_grad(t_scale, grad=False)
        m.train(300, 1024, 3e-3, t_scale, seed=1)
        l1 = m.loss_grad(t_scale, grad=False)
        out = m.forward()
        th1 = m.get_params()
    print("loss", l0, "->", l1, "nu_t median", np.median(out[0]))
    assert l1 < 0.05 * l0
    assert abs(np.median(out[0]) / nu_star - 1) < 0.2
    assert np.array_equal(th1[9479:], th[9479:])   # the scales stay


def _smooth_state(n, seed, ncomp):
    rng = np.random.default_rng(seed)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    t = 2 * np.pi / n
    rho = 1 + 1e-3 * np.sin(t * (i + 2 * k))
    u = 0.03 * np.sin(t * j) * np.cos(t * k) + 2e-3 * rng.standard_normal((n, n, n))
    v = 0.02 * np.cos(t * k) + 2e-3 * rng.standard_normal((n, n, n))
    w = 0.01 * np.sin(t * i) * np.sin(t * j)
    r = np.sqrt((i - n / 2) ** 2 + (j - n / 2) ** 2 + (k - n / 2) ** 2)
    phi = 0.5 * (1 + np.tanh(0.5 * (n / 4 - r))) if ncomp == 2 else np.zeros_like(u)
    return [a.astype(np.float32) for a in (rho, u, v, w, phi)]


@pytest.mark.parametrize("ncomp,r", [(1, 1), (1, 4), (2, 2), (2, 4)])
def test_samples_match_oracle_filter_and_features(ncomp, r):
    n = 32
    fields = _smooth_state(n, 8, ncomp)
    with Lattice(n, n, n, ncomp=ncomp, tau=0.7, sigma=0.01 if ncomp == 2 else 0.0,
                 sharp_c=2.0 if ncomp == 2 else 0.0) as lat, Closure(max_samples=2 * (n // r) ** 3) as m:
        for name, a in zip(("rho", "u", "v", "w", "phi"), fields):
            lat.upload_field(name, a)
        m.add_samples(lat, r)
        m.add_samples(lat, r)
        assert m.num_samples == 2 * (n // r) ** 3
        feat, aux = m.get_samples()
    f_ref, a_ref = cl.samples(*[a.astype(np.float64) for a in fields], r)
    nc = (n // r) ** 3
    # per group (strain, grad phi, phi): rounding relative to the group's size
    groups = [range(0, 6)] * 6 + [range(6, 9)] * 3 + [range(9, 10)]
    for half in (slice(0, nc), slice(nc, 2 * nc)):
        for q in range(10):
            scale = max(np.abs(f_ref[:, list(groups[q])]).max(), 1e-12)
            assert np.abs(feat[q, half] - f_ref[:, q]).max() <= 2e-5 * scale, (q, r)
        assert np.abs(aux[0, half] - a_ref[:, 0]).max() < 1e-6
        tscale = np.abs(a_ref[:, 1:]).max()
        if r > 1:
            assert np.abs(aux[1:, half].T - a_ref[:, 1:]).max() <= 1e-4 * tscale
        else:
            assert np.abs(aux[1:, half]).max() == 0.0


def _random_closure_fields(n, seed):
    rng = np.random.default_rng(seed)
    clo = np.zeros((7, n, n, n), np.float32)
    clo[0] = 0.02 * rng.uniform(size=(n, n, n))
    clo[1:] = 2e-5 * rng.standard_normal((6, n, n, n))
    return clo


@pytest.mark.parametrize("ncomp,storage_walls", [(1, False), (1, True), (2, False)])
def test_collision_with_given_closure_fields_matches_oracle(ncomp, storage_walls):
    n, steps = 24, 10
    fields = _smooth_state(n, 9, ncomp)
    clo = _random_closure_fields(n, 10)
    kw = dict(ncomp=ncomp, tau=0.6, sigma=0.01 if ncomp == 2 else 0.0, sharp_c=2.0 if ncomp == 2 else 0.0,
              periodic=(1, 0, 1) if storage_walls else (1, 1, 1))
    isf = None
    if storage_walls:
        isf = np.ones((n, n, n), np.int32)
        isf[8:12, 5:9, 10:14] = 0
    with Lattice(n, n, n, **kw) as lat:
        if isf is not None:
            lat.set_isfluid(isf)
        for name, a in zip(("rho", "u", "v", "w", "phi"), fields):
            lat.upload_field(name, a)
        lat.init_equilibrium(True)
        lat.set_closure_fields(clo)
        assert np.array_equal(lat.download_closure(), clo)
        lat.step(steps)
        got = lat.download()
    o = po.Oracle(n, n, n, **kw)
    if isf is not None:
        o.set_isfluid(isf)
    o.set_closure(clo.astype(np.float64))
    o.init(*[a.astype(np.float64) for a in fields], prepush=True)
    o.step(steps)
    ref = o.moments()
    o.close()
    err = field_errors(got, ref)
    print({k: f"{v:.2e}" for k, v in err.items()})
    for f, v in err.items():
        assert v <= 1e-5, (f, v)


@pytest.mark.parametrize("ncomp", [1, 2])
def test_attached_model_matches_oracle_loop(ncomp):
    """A model attached to a lattice: before every collision the inputs of
    the current state go through the MLP; the oracle repeats that per step."""
    n, steps = 16, 4
    fields = _smooth_state(n, 11, ncomp)
    th = cl.init_params(12, in_scale=np.array([50.0] * 6 + [10.0] * 3 + [2.0]),
                        out_scale=np.array([5e-3] + [2e-5] * 6))
    kw = dict(ncomp=ncomp, tau=0.6, sigma=0.01 if ncomp == 2 else 0.0, sharp_c=2.0 if ncomp == 2 else 0.0)
    with Lattice(n, n, n, **kw) as lat, Closure() as m:
        m.set_params(th)
        for name, a in zip(("rho", "u", "v", "w", "phi"), fields):
            lat.upload_field(name, a)
        lat.init_equilibrium(True)
        lat.set_closure(m)
        clo0 = lat.download_closure()
        lat.step(steps)
        got = lat.download()
        with pytest.raises(TslbError):
            lat.set_jet_bc(True)
    o = po.Oracle(n, n, n, **kw)
    o.init(*[a.astype(np.float64) for a in fields], prepush=True)
    for s in range(steps):
        mm = o.moments()
        f = cl.features(mm[1], mm[2], mm[3], mm[10] if ncomp == 2 else 0 * mm[1]).reshape(-1, 10)
        out, _ = cl.forward(th, f.astype(np.float32), emulate=True)
        clo = out.T.reshape(7, n, n, n)
        if s == 0:
            assert np.abs(clo0 - clo).max() <= 2e-3 * np.abs(clo).max()
        o.set_closure(clo)
        o.step(1)
    ref = o.moments()
    o.close()
    err = field_errors(got, ref)
    print({k: f"{v:.2e}" for k, v in err.items()})
    for fname, v in err.items():
        assert v <= 1e-4, (fname, v)


def test_closure_argument_checks():
    with Closure(max_samples=10) as m:
        with pytest.raises(ValueError):
            m.set_params(np.zeros(5, np.float32))
        with pytest.raises(TslbError):
            m.train(1, 8, 1e-3, 1.0)          # no samples
        with Lattice(8, 8, 8, ncomp=1, tau=0.7) as lat:
            with pytest.raises(TslbError):
                m.add_samples(lat, 3)         # 3 does not divide 8
            with pytest.raises(TslbError):
                m.add_samples(lat, 1)         # 512 samples > capacity 10
            with pytest.raises(TslbError):
                lat.download_closure()        # nothing attached
    with Lattice(8, 8, 8, ncomp=3, tau=0.7) as lat, Closure() as m:
        with pytest.raises(TslbError):
            lat.set_closure(m)                # colour model: not supported
