"""Pins for the LES-with-learned-closure oracle (oracle/les_cnn.py): central
finite differences (eps = 1e-3) of the FP64 rollout loss along 8 random
parameter directions and along a random initial-condition direction, a
dot-product test of the single RHS transpose, and the closure-free case
against the DNS adjoint."""
import numpy as np
import pytest
import torch

from oracle import adjoint as A
from oracle import les_cnn as L
from oracle import spectral as S

from test_adjoint_oracle import dot, field


def small_net(seed, width=8, s_in=1.0, s_out=0.2, extra=0.3):
    r = np.random.default_rng(seed)
    shapes = [(width, 4), (width, width), (width, width), (6, width)]
    Ws = [r.standard_normal((27, co, ci)) * np.sqrt(2.0 / (27 * (ci + co))) for co, ci in shapes]
    bs = [r.standard_normal(width) * 0.1 for _ in range(3)]
    return L.Net(Ws, bs, 4, s_in, s_out, extra, emulate=False, dtype=torch.float64)


def setup(n=12, seed=0):
    g = S.Grid(3, n)
    u0 = field(g, seed, scale=0.05)
    refs = [field(g, seed + 10 + t, scale=0.05) for t in range(2)]
    return g, u0, refs


def test_closure_is_zero_for_zero_output_layer():
    g, u0, _ = setup()
    net = small_net(1)
    net.Ws[3][:] = 0
    assert np.allclose(L.les_rhs(u0, g, 0.05, net), S.convectiondiffusion(u0, g, 0.05), atol=1e-14)


def test_les_rhs_vjp_dot_product():
    g, u, _ = setup(seed=3)
    net = small_net(2)
    du = field(g, 7)
    w = field(g, 8, trunc=False)
    nu, eps = 0.05, 1e-4
    f = lambda v: L.les_rhs(v, g, nu, net)
    Jdu = (f(u + eps * du) - f(u - eps * du)) / (2 * eps)
    assert dot(w, Jdu, g) == pytest.approx(dot(L.les_rhs_vjp(u, w, g, nu, net), du, g), rel=1e-6)


def test_rollout_gradient_finite_differences():
    """the spec's check: eps = 1e-3 on 8 random unit parameter directions,
    plus one initial-condition direction"""
    g, u0, refs = setup(seed=5)
    net = small_net(4)
    nu, dt = 0.05, 0.05
    Lval, ubar, (gW, gb) = L.rollout_gradient(u0, refs, dt, g, nu, net)
    assert Lval > 0
    r = np.random.default_rng(99)
    eps = 1e-3
    for _ in range(8):
        dW = [r.standard_normal(W.shape) for W in net.Ws]
        db = [r.standard_normal(np.shape(b)) for b in net.bs]
        nrm = np.sqrt(sum(np.sum(d * d) for d in dW) + sum(np.sum(d * d) for d in db))
        dW = [d / nrm for d in dW]
        db = [d / nrm for d in db]
        lp, _ = L.rollout_loss(u0, refs, dt, g, nu, net.perturbed(dW, db, eps))
        lm, _ = L.rollout_loss(u0, refs, dt, g, nu, net.perturbed(dW, db, -eps))
        fd = (lp - lm) / (2 * eps)
        an = sum(np.sum(a * d) for a, d in zip(gW, dW)) + sum(np.sum(a * d) for a, d in zip(gb, db))
        assert fd == pytest.approx(an, rel=1e-5, abs=1e-9 * abs(Lval)), (fd, an)
    du = field(g, 77)
    lp, _ = L.rollout_loss(u0 + eps * du, refs, dt, g, nu, net)
    lm, _ = L.rollout_loss(u0 - eps * du, refs, dt, g, nu, net)
    assert (lp - lm) / (2 * eps) == pytest.approx(dot(ubar, du, g), rel=1e-5)


def test_step_vjp_without_closure_matches_dns_adjoint():
    g, u, _ = setup(seed=6)
    net = small_net(5)
    net.Ws[3][:] = 0
    ub = field(g, 9, trunc=False)
    got = L.rk4_vjp(u, ub, 0.05, g, 0.05, net)
    ref = A.rk4_vjp(u, ub, 0.05, g, 0.05)
    assert np.linalg.norm(got - ref) < 1e-12 * np.linalg.norm(ref)
