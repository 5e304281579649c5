"""Pins for the training-semantics oracle (oracle/optim.py, the trace removal
and relative loss of oracle/cnn.py): closed forms and special cases, chosen
so that a dropped bias correction, a global instead of per-array clip, an
uncoupled decay, an off-by-one in the schedule or a missing projection in the
backward fails one of them."""
import math

import numpy as np
import torch

from helpers import oracle_solenoidal
from oracle import cnn as OC
from oracle import optim as O
from oracle import spectral as S

LR = 1e-3


def test_adam_first_step_is_sign_step():
    g = np.array([3.0, -1e-3, 250.0, 0.0])
    for scale in (1.0, 1e-4, 1e4):
        st = O.AdamWState(4)
        w = O.clipnorm_adamw(np.zeros(4), scale * g, st, [0, 4], LR)
        ref = -LR * scale * g / (np.abs(scale * g) + 1e-8)
        assert np.allclose(w, ref, rtol=1e-12, atol=0)


def test_clip_two_steps_equal_unit_gradient_steps():
    d = np.array([0.6, -0.8, 0.0])  # unit norm
    bounds = [0, 3]
    st_c, st_u = O.AdamWState(3), O.AdamWState(3)
    w_c = w_u = np.zeros(3)
    for s in (10.0, 2.0):
        w_c = O.clipnorm_adamw(w_c, s * d, st_c, bounds, LR, clip=1.0)
        w_u = O.clipnorm_adamw(w_u, d, st_u, bounds, LR)
    assert np.allclose(w_c, w_u, rtol=1e-14, atol=0)
    # without the clip the same gradients give a different second step:
    # m1 = 0.1*10d = d, m2 = 0.9 d + 0.1*2d = 1.1 d, mhat2 = 1.1 d / 0.19;
    # v2 = 0.999*0.001*100 d^2 + 0.001*4 d^2 = 0.1039 d^2, vhat2 = v2 / (1 - 0.999^2)
    st = O.AdamWState(3)
    w = O.clipnorm_adamw(np.zeros(3), 10 * d, st, bounds, LR)
    w = O.clipnorm_adamw(w, 2 * d, st, bounds, LR)
    mhat = 1.1 / 0.19
    vhat = 0.1039 / (1 - 0.999**2)
    step2 = LR * mhat * np.abs(d) / (math.sqrt(vhat) * np.abs(d) + 1e-8)
    exp = -(LR * np.sign(d) * (np.abs(d) > 0)) - np.sign(d) * step2 * (np.abs(d) > 0)
    assert np.allclose(w, exp, rtol=1e-9, atol=1e-15)
    assert not np.allclose(w, w_c, rtol=1e-3)


def test_clip_is_per_array():
    """array 0 above the clip, array 1 below: array 1's steps are those of
    the unclipped optimiser; array 0's those of its unit-norm gradient"""
    big, small = np.array([3.0, 4.0]), np.array([0.3, -0.4])
    bounds = [0, 2, 4]
    st = O.AdamWState(4)
    w = np.zeros(4)
    st0, st1 = O.AdamWState(2), O.AdamWState(2)
    w0, w1 = np.zeros(2), np.zeros(2)
    for s in (1.0, 3.0):
        w = O.clipnorm_adamw(w, np.concatenate([s * big, s * small]), st, bounds, LR, clip=1.0)
        w0 = O.clipnorm_adamw(w0, big / 5.0, st0, [0, 2], LR)
        g1 = s * small
        g1 = g1 * min(1.0, 1.0 / np.linalg.norm(g1))
        w1 = O.clipnorm_adamw(w1, g1, st1, [0, 2], LR)
    assert np.allclose(w[:2], w0, rtol=1e-14)
    assert np.allclose(w[2:], w1, rtol=1e-14)
    # a single global clip would scale array 1 too (its norm after the second
    # scaling, 1.5, is above 1): the per-array result differs from it
    stg = O.AdamWState(4)
    wg = np.zeros(4)
    for s in (1.0, 3.0):
        gg = np.concatenate([s * big, s * small])
        wg = O.clipnorm_adamw(wg, gg * min(1.0, 1.0 / np.linalg.norm(gg)), stg, [0, 4], LR)
    assert not np.allclose(wg[2:], w[2:], rtol=1e-6)


def test_decay_alone():
    w0 = np.array([1.0, -2.0, 0.5])
    w = O.clipnorm_adamw(w0, np.zeros(3), O.AdamWState(3), [0, 3], LR, weight_decay=0.1)
    assert np.allclose(w, w0 * (1 - LR * 0.1), rtol=1e-15)


def test_schedule_end_points():
    total = 200  # nwarm = 10
    assert O.lr_schedule(1, total, LR) == LR / 10
    assert O.lr_schedule(9, total, LR) == LR * 9 / 10
    assert O.lr_schedule(10, total, LR) == LR
    assert abs(O.lr_schedule(105, total, LR) - LR / 2) < 1e-18
    assert O.lr_schedule(total, total, LR) == 0.0
    assert O.lr_schedule(1, 5, LR) == LR  # nwarm = max(1, 0)


def test_relative_loss_special_cases_and_gradient():
    r = np.random.default_rng(0)
    t = torch.tensor(r.standard_normal((2, 64, 6)), dtype=torch.float32)
    y0 = torch.zeros(2, 64, 16)
    l0, _ = OC.rel_loss(y0, t, 6)
    assert abs(l0 - 1) < 1e-6
    y2 = torch.zeros(2, 64, 16)
    y2[..., :6] = 2 * t
    l2, _ = OC.rel_loss(y2, t, 6)
    assert abs(l2 - 1) < 1e-6
    y = torch.tensor(r.standard_normal((2, 64, 16)), dtype=torch.float32)
    l, dy = OC.rel_loss(y, t, 6)
    assert torch.all(dy[..., 6:] == 0)
    for j in [(0, 3, 1), (1, 60, 5), (0, 10, 0)]:
        h = 1e-2
        yp, ym = y.clone(), y.clone()
        yp[j] += h
        ym[j] -= h
        fd = (OC.rel_loss(yp, t, 6)[0] - OC.rel_loss(ym, t, 6)[0]) / (2 * h)
        assert abs(fd - float(dy[j])) < 1e-6 * max(1.0, abs(fd))
    # the MSE rescaled by N / sum t^2
    lm, _ = OC.mse(y, t, 6)
    N = t.numel()
    assert abs(l - lm * N / (float((t.double() ** 2).sum()) + 1.1920929e-7)) < 1e-9


def test_tracefree_projection_and_backward_against_autograd():
    """the projected output has zero trace and is unchanged by a second
    projection; the explicit backward through it equals autograd"""
    from test_cnn_oracle import params

    n, B = 5, 1
    Ws, bs = params(7)
    Ws[3] = np.random.default_rng(8).standard_normal((27, 6, 8)).astype(np.float32) * 0.2
    x = torch.tensor(np.random.default_rng(9).standard_normal((B, n**3, 4)), dtype=torch.float32)
    y, store = OC.forward(Ws, bs, x, n, emulate=False, tracefree_out=True)
    assert float(y[..., :3].sum(-1).abs().max()) < 1e-5
    assert torch.allclose(OC.tracefree(y, dim=2), y, atol=1e-6)
    y_raw, _ = OC.forward(Ws, bs, x, n, emulate=False)
    assert torch.allclose(y[..., 3:], y_raw[..., 3:])
    Wt = [OC.torch_weight(W).requires_grad_(True) for W in Ws]
    bt = [torch.tensor(b).requires_grad_(True) for b in bs]
    xs = OC.to_ncdhw(x, n).requires_grad_(True)
    a = xs
    for l in range(3):
        a = OC.gelu(OC.conv(a, Wt[l]) + bt[l].view(1, -1, 1, 1, 1))
    z = OC.conv(a, Wt[3])
    tr = (z[:, 0] + z[:, 1] + z[:, 2]) / 3
    yy = torch.cat([z[:, :3] - tr.unsqueeze(1), z[:, 3:]], dim=1)
    tgt = torch.randn(B, n**3, 6, generator=torch.Generator().manual_seed(3))
    loss = ((OC.to_nvc(yy) - tgt) ** 2).sum()
    loss.backward()
    dy = 2 * (y - tgt)
    gW, gb, dx = OC.backward(store, dy, n, emulate=False)
    rel = lambda p, q: float(torch.linalg.norm(p - q) / torch.linalg.norm(q))
    for l in range(4):
        ref = Wt[l].grad.permute(2, 3, 4, 0, 1).reshape(27, *Wt[l].shape[:2])
        assert rel(gW[l], ref) < 1e-5, l
    for l in range(3):
        assert rel(gb[l], bt[l].grad) < 1e-5
    assert rel(dx, OC.to_nvc(xs.grad)) < 1e-5


def test_e_post_special_cases():
    g = S.Grid(3, 16)
    u = oracle_solenoidal(g, 5)
    tr = lambda c: S.twothirds(c, g)
    ut = np.stack([tr(c) for c in u])
    assert O.e_post(ut, u, tr) == 0.0
    assert abs(O.e_post(np.zeros_like(u), u, tr) - 1) < 1e-15
    assert abs(O.e_post(2 * ut, u, tr) - 1) < 1e-15
    assert abs(O.e_post(0.5 * ut, u, tr) - 0.5) < 1e-15
