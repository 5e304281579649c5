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


def rand_params(D, widths, use_redelta, seed, tbnn=False):
    rng = np.random.default_rng(seed)
    shapes = OP.layer_shapes(D, widths, use_redelta, tbnn)
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


# ------------------------------------------------------------------ TBNN
def rand_orthogonal(D, seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((D, D)))
    return q


def rotate_grad(GG, Q, D):
    """component-major gradient of Q A Q^T (A[i][j] at component i + D j)"""
    A, _ = OP.grad_matrices(GG, D)
    B = Q @ A @ Q.T
    shape = GG.shape[1:]
    return np.stack([B[:, i, j].reshape(shape) for j in range(D) for i in range(D)])


def unpack(bas, D):
    """packed [.., nt] -> [.., D, D]"""
    out = np.zeros(bas.shape[:-1] + (D, D))
    comps = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0)] if D == 3 else [(0, 0), (1, 1), (0, 1)]
    for t, (a, b) in enumerate(comps):
        out[..., a, b] = bas[..., t]
        out[..., b, a] = bas[..., t]
    return out


@pytest.mark.parametrize("D", [2, 3])
def test_tb_invariance_and_equivariance(D):
    GG = np.random.default_rng(3).standard_normal((D * D, 5, 5))
    Q = rand_orthogonal(D, 4)
    inv, bas, a2 = OP.tb_features(GG, D)
    inv2, bas2, a22 = OP.tb_features(rotate_grad(GG, Q, D), D)
    assert np.allclose(a22, a2, rtol=1e-13)
    assert np.allclose(inv2, inv, rtol=1e-11, atol=1e-13)
    B = unpack(bas, D)
    B2 = unpack(bas2, D)
    assert np.allclose(B2, Q @ B @ Q.T, atol=1e-12)
    # symmetric (by packing) and trace-free
    assert np.abs(np.trace(B, axis1=-2, axis2=-1)).max() < 1e-14


def test_tb_closed_forms():
    """pure strain A = diag(1, 2, -3) and pure rotation A = [[0, w], [-w, 0]]"""
    GG = np.zeros((9, 1))
    for i, v in enumerate((1.0, 2.0, -3.0)):
        GG[i + 3 * i, 0] = v
    inv, bas, a2 = OP.tb_features(GG, 3)
    n = math.sqrt(14.0) + np.finfo(float).eps
    s = np.array([1.0, 2.0, -3.0]) / n
    assert a2[0] == 14.0
    assert inv[0] == pytest.approx([np.sum(s**2), 0, np.sum(s**3), 0, 0], abs=1e-15)
    assert np.allclose(bas[0, 0, :3], s - s.mean(), atol=1e-15) and np.all(bas[0, 0, 3:] == 0)
    assert np.allclose(bas[0, 1, :3], s**2 - np.mean(s**2), atol=1e-15)
    assert np.abs(bas[0, 2:]).max() < 1e-15  # every R-term vanishes
    # 2D pure rotation: S = 0, tr(R R) = -2 r^2 with r = 1/sqrt(2)
    G2 = np.array([[0.0], [-1.0], [1.0], [0.0]])  # xx, yx, xy, yy
    inv, bas, _ = OP.tb_features(G2, 2)
    assert inv[0] == pytest.approx([0.0, -1.0], abs=1e-15)
    assert np.abs(bas).max() == 0


@pytest.mark.parametrize("D,use_redelta", [(3, False), (3, True), (2, True)])
def test_tbnn_forward_backward_match_autograd(D, use_redelta):
    p, shapes = rand_params(D, (6, 5, 4), use_redelta, 8, tbnn=True)
    nin = shapes[0][0]
    GG = np.random.default_rng(9).standard_normal((D * D, 30))
    x = OP.tb_rows(GG, D, z=0.3 if use_redelta else None)
    m, cache = OP.tbnn_forward(p, x, D, nin)
    tp = {k: torch.tensor(v, requires_grad=True) for k, v in p.items()}
    a = torch.tensor(x[:, :nin])
    for l in range(1, 4):
        a = torch.nn.functional.gelu(a @ tp[f"W{l}"] + tp[f"b{l}"], approximate="tanh")
    w = a @ tp["W4"]
    nt, nb = (6, 7) if D == 3 else (3, 2)
    B = torch.tensor(x[:, nin:]).reshape(-1, nb, nt)
    tm = (w[:, :, None] * B).sum(dim=1)
    assert np.allclose(m, tm.detach().numpy(), rtol=1e-13, atol=1e-14)
    target = np.random.default_rng(10).standard_normal(m.shape)
    loss, dm = OP.relative_loss(m, target)
    tl = ((tm - torch.tensor(target)) ** 2).sum() / ((torch.tensor(target) ** 2).sum() + np.finfo(np.float32).eps)
    tl.backward()
    g = OP.tbnn_backward(p, cache, dm)
    for k in p:
        assert np.allclose(g[k], tp[k].grad.numpy(), rtol=1e-11, atol=1e-14), k
    # the TBNN closure is equivariant under any rotation (z is invariant)
    re = (1e-3, 0.5, 2.0) if use_redelta else None
    tau = OP.tbnn_model(p, GG.reshape(D * D, 5, 6), 0.3, D, redelta=re)
    Q = rand_orthogonal(D, 11)
    tau2 = OP.tbnn_model(p, rotate_grad(GG.reshape(D * D, 5, 6), Q, D), 0.3, D, redelta=re)
    T1 = unpack(np.moveaxis(tau, 0, -1), D)
    T2 = unpack(np.moveaxis(tau2, 0, -1), D)
    assert np.allclose(T2, Q @ T1 @ Q.T, atol=1e-10 * np.abs(T1).max())


# ------------------------------------------------- the cube's symmetry group
@pytest.mark.parametrize("D", [2, 3])
def test_group_tables(D):
    els = OP.group_elements(D)
    mats = [OP.element_matrix(p, s) for p, s in els]
    assert len(mats) == (48 if D == 3 else 8)
    keys = {m.tobytes() for m in mats}
    assert len(keys) == len(mats)
    assert np.array_equal(mats[0], np.eye(D))
    for a in mats:
        assert np.allclose(a @ a.T, np.eye(D))
        for b in mats[:8]:
            assert (a @ b).tobytes() in keys
    # (R A R^T)_ab = s_a s_b A_{p_a p_b}
    A = np.random.default_rng(0).standard_normal((D, D))
    p, s = els[-1]
    R = OP.element_matrix(p, s)
    RA = R @ A @ R.T
    for a in range(D):
        for b in range(D):
            assert RA[a, b] == pytest.approx(s[a] * s[b] * A[p[a], p[b]], abs=1e-15)


@pytest.mark.parametrize("D", [2, 3])
def test_symmetrize_equivariant_and_idempotent(D):
    p, _ = rand_params(D, (5, 6, 7), False, 13)
    GG = np.random.default_rng(14).standard_normal((D * D, 4, 5))
    model = lambda G: OP.fullchain(p, G, 0.3, D)
    msym = OP.symmetrize(model, GG, D)
    for pe, se in OP.group_elements(D)[1:]:
        R = OP.element_matrix(pe, se)
        a = unpack(np.moveaxis(OP.symmetrize(model, rotate_grad(GG, R, D), D), 0, -1), D)
        b = unpack(np.moveaxis(msym, 0, -1), D)
        assert np.allclose(a, R @ b @ R.T, atol=1e-13 * np.abs(b).max())
    # the plain mlp is not equivariant, so the average changes it ...
    assert not np.allclose(msym, model(GG), atol=1e-6)
    # ... while an equivariant model is its own average
    pt, _ = rand_params(D, (5, 6, 7), False, 15, tbnn=True)
    tb = lambda G: OP.tbnn_model(pt, G, 0.3, D)
    assert np.allclose(OP.symmetrize(tb, GG, D), tb(GG), atol=1e-13 * np.abs(tb(GG)).max())


# ------------------------------------------------------ equivariant_net
def test_equi_synthesis_in_reynolds_fixed_space_2d():
    """the synthesis columns are orthonormal fixed points of the Reynolds
    projectors r / |G| of get_weight_projectors (src/symmetry.jl:92-127),
    built here by brute force for the 8-element 2D group"""
    D, nten = 2, 4
    els = OP.group_elements(D)
    cayley, _ = OP.cayley_tables(D)
    n = len(els)
    s_lift, s_mid, s_sink = OP.weight_synthesis(D)
    r_lift = np.zeros((n * nten, n * nten))
    r_sink = np.zeros((nten * n, nten * n))
    for m in range(n):
        for x in range(D):
            for y in range(D):
                for nn in range(n):
                    for i in range(D):
                        for j in range(D):
                            v = sum(s[x] * s[y] * (cayley[g, nn] == m) * (p[x] == i) * (p[y] == j)
                                    for g, (p, s) in enumerate(els))
                            r_lift[m + n * (x + D * y), nn + n * (i + D * j)] = v
                            r_sink[x + D * y + nten * m, i + D * j + nten * nn] = v
    r_mid = np.zeros((n * n, n * n))
    for m in range(n):
        for nn in range(n):
            for a in range(n):
                for b in range(n):
                    r_mid[m + n * nn, a + n * b] = sum((cayley[g, a] == m) * (cayley[g, b] == nn) for g in range(n))
    for r, sy in ((r_lift, s_lift), (r_mid, s_mid), (r_sink, s_sink)):
        assert np.allclose(r / n @ sy, sy, atol=1e-12)
        assert np.allclose(sy.T @ sy, np.eye(sy.shape[1]), atol=1e-6)
    # the sum of the projector's fixed-space dimension equals the basis size
    assert round(np.trace(r_mid / n)) == s_mid.shape[1]


def rand_equi(D, widths, use_redelta, seed):
    sh = OP.equi_shapes(D, widths, use_redelta)
    rng = np.random.default_rng(seed)
    return np.concatenate([0.5 * rng.standard_normal(int(np.prod(s))) for _, s in sh])


@pytest.mark.parametrize("D,use_redelta", [(2, False), (3, True)])
def test_equi_network_equivariant(D, use_redelta):
    widths = (2, 3, 2)
    h = rand_equi(D, widths, use_redelta, 5)
    layers = OP.equi_project(OP.equi_unflat(h, D, widths, use_redelta), D, OP.weight_synthesis(D))
    GG = np.random.default_rng(6).standard_normal((D * D, 7))
    x, _ = OP.features(GG, None, 0.3, 0.4 if use_redelta else None)
    y = OP.equi_net(layers, torch.tensor(x), D).numpy()
    Y = unpack(y, D)
    assert np.abs(np.trace(Y, axis1=-2, axis2=-1)).max() < 1e-12
    for p, s in OP.group_elements(D)[1:]:
        R = OP.element_matrix(p, s)
        xr, _ = OP.features(rotate_grad(GG, R, D), None, 0.3, 0.4 if use_redelta else None)
        yr = unpack(OP.equi_net(layers, torch.tensor(xr), D).numpy(), D)
        assert np.allclose(yr, R @ Y @ R.T, atol=1e-12 * np.abs(Y).max())
    # a plain mlp of the same widths is not equivariant (the check can fail)
    p_, _ = rand_params(D, widths, use_redelta, 7)
    yp = unpack(OP.forward(p_, x, D)[0], D)
    pe, se = OP.group_elements(D)[1]
    R = OP.element_matrix(pe, se)
    xr, _ = OP.features(rotate_grad(GG, R, D), None, 0.3, 0.4 if use_redelta else None)
    assert not np.allclose(unpack(OP.forward(p_, xr, D)[0], D), R @ yp @ R.T, atol=1e-6)
