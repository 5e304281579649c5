"""Pins for the pointwise MLP oracle (oracle/pmlp.py): forward and every
parameter gradient against torch.autograd of the same network built from
torch.nn.functional (an independent formulation: a dropped GELU' factor, a
transposed weight, a missing trace removal in the backward or a bias gradient
summed over the wrong axis fails it), the trace removal's invariants, the
degree-two homogeneity of fullchain and the Re_Delta input's shift, and the
features' normalisation."""
import math

import numpy as np
import pytest
import torch

from oracle import pmlp as OP


def rand_params(D, widths, use_redelta, seed):
    rng = np.random.default_rng(seed)
    shapes = OP.layer_shapes(D, widths, use_redelta)
    p = {}
    for l, (ci, co) in enumerate(shapes):
        p[f"W{l + 1}"] = rng.uniform(-1, 1, (ci, co)) * math.sqrt(6 / (ci + co))
        if l < 3:
            p[f"b{l + 1}"] = 0.1 * rng.standard_normal(co)
    return p, shapes


def torch_net(p, x, D):
    """the network in torch (float64) with autograd leaves"""
    tp = {k: torch.tensor(v, requires_grad=True) for k, v in p.items()}
    a = torch.tensor(x)
    for l in range(1, 4):
        a = torch.nn.functional.gelu(a @ tp[f"W{l}"] + tp[f"b{l}"], approximate="tanh")
    y = a @ tp["W4"]
    k = 3 if D == 3 else 2
    y = torch.cat([y[:, :k] - y[:, :k].mean(dim=1, keepdim=True), y[:, k:]], dim=1)
    return y, tp


@pytest.mark.parametrize("D,use_redelta", [(3, False), (3, True), (2, False)])
def test_forward_backward_match_autograd(D, use_redelta):
    p, shapes = rand_params(D, (7, 5, 6), use_redelta, 3)
    nin = shapes[0][0]
    x = np.random.default_rng(4).standard_normal((40, nin))
    y, cache = OP.forward(p, x, D)
    ty, tp = torch_net(p, x, D)
    assert np.allclose(y, ty.detach().numpy(), rtol=1e-13, atol=1e-14)
    target = np.random.default_rng(5).standard_normal(y.shape)
    loss, dy = OP.relative_loss(y, target)
    tl = ((ty - torch.tensor(target)) ** 2).sum() / ((torch.tensor(target) ** 2).sum() + np.finfo(np.float32).eps)
    tl.backward()
    assert loss == pytest.approx(tl.item(), rel=1e-13)
    g = OP.backward(p, cache, dy, D)
    for k in p:
        assert np.allclose(g[k], tp[k].grad.numpy(), rtol=1e-11, atol=1e-14), k


def test_mse_loss_matches_torch():
    a = np.random.default_rng(1).standard_normal((30, 6))
    b = np.random.default_rng(2).standard_normal((30, 6))
    l, d = OP.mse_loss(a, b)
    ta = torch.tensor(a, requires_grad=True)
    tl = torch.nn.functional.mse_loss(ta, torch.tensor(b))
    tl.backward()
    assert l == pytest.approx(tl.item(), rel=1e-14)
    assert np.allclose(d, ta.grad.numpy(), rtol=1e-13)


def test_gelu_values():
    assert OP.gelu(np.array([0.0]))[0] == 0.0
    assert OP.gelu(np.array([10.0]))[0] == pytest.approx(10.0, rel=1e-15)
    assert abs(OP.gelu(np.array([-10.0]))[0]) < 1e-15
    z = np.linspace(-4, 4, 101)
    h = 1e-6
    assert np.allclose(OP.gelu_grad(z), (OP.gelu(z + h) - OP.gelu(z - h)) / (2 * h), atol=1e-9)


@pytest.mark.parametrize("D", [2, 3])
def test_tracefree_invariants(D):
    y = np.random.default_rng(0).standard_normal((20, 6 if D == 3 else 3))
    t = OP.tracefree(y, D)
    k = 3 if D == 3 else 2
    assert np.abs(t[:, :k].sum(axis=1)).max() < 1e-14
    assert np.array_equal(t[:, k:], y[:, k:])
    assert np.allclose(OP.tracefree(t, D), t, atol=1e-15)


def test_flat_roundtrip_and_bounds():
    p, shapes = rand_params(3, (4, 5, 6), True, 2)
    h = OP.flat(p, shapes)
    b = OP.array_bounds(shapes)
    assert b[-1] == h.size == 10 * 4 + 4 + 4 * 5 + 5 + 5 * 6 + 6 + 6 * 6
    q = OP.unflat(h, shapes)
    for k in p:
        assert np.array_equal(p[k], q[k])
    # W1 column-major (c_in fastest): element (ci=1, co=0) is the second entry
    assert h[1] == p["W1"][1, 0]


def test_features_normalisation():
    rng = np.random.default_rng(7)
    GG = rng.standard_normal((9, 4, 4, 4))
    tau = rng.standard_normal((6, 4, 4, 4))
    delta = 0.3
    x, y = OP.features(GG, tau, delta, z=0.25)
    assert np.allclose(np.linalg.norm(x[:, :9], axis=1), 1.0, atol=1e-14)
    assert np.all(x[:, 9] == 0.25)
    a2 = (GG**2).sum(axis=0).ravel()
    assert np.allclose(y * (delta**2 * a2)[:, None], tau.reshape(6, -1).T, rtol=1e-13)
    # eps matters where the gradient vanishes
    GG[:, 0, 0, 0] = 0
    x, y = OP.features(GG, tau, delta)
    assert np.all(x[0] == 0) and np.allclose(y[0], tau[:, 0, 0, 0] / np.finfo(float).eps)


@pytest.mark.parametrize("D", [2, 3])
def test_fullchain_homogeneous_of_degree_two(D):
    p, _ = rand_params(D, (5, 6, 7), False, 11)
    shp = (D * D,) + (6,) * D
    GG = np.random.default_rng(8).standard_normal(shp)
    t1 = OP.fullchain(p, GG, 0.4, D)
    t2 = OP.fullchain(p, 3.0 * GG, 0.4, D)
    assert np.allclose(t2, 9.0 * t1, rtol=1e-12, atol=1e-14)
    # the trace of the diagonal components vanishes pointwise
    k = 3 if D == 3 else 2
    assert np.abs(t1[:k].sum(axis=0)).max() < 1e-12 * np.abs(t1).max()


def test_fullchain_redelta_shift():
    """scaling G by a multiplies Re_Delta by a, so z moves by log(a) / sigma:
    the network sees the same normalised gradient and a shifted z"""
    p, _ = rand_params(3, (5, 6, 7), True, 12)
    GG = np.random.default_rng(9).standard_normal((9, 5, 5, 5))
    visc, mu, sigma, delta, a = 1e-3, 1.2, 0.7, 0.3, 2.5
    z1 = OP.redelta_z(GG, delta, visc, mu, sigma)
    z2 = OP.redelta_z(a * GG, delta, visc, mu, sigma)
    assert z2 - z1 == pytest.approx(math.log(a) / sigma, rel=1e-12)
    re = delta**2 * math.sqrt((GG**2).sum(axis=0).mean()) / visc
    assert z1 == pytest.approx((math.log(re) - mu) / sigma, rel=1e-14)
    # with mu moved by log(a) the network sees the same z, so tau(aG) = a^2 tau(G)
    t1 = OP.fullchain(p, GG, delta, 3, redelta=(visc, mu, sigma))
    t2 = OP.fullchain(p, a * GG, delta, 3, redelta=(visc, mu + math.log(a), sigma))
    assert np.allclose(t2, a * a * t1, rtol=1e-12, atol=1e-14)
