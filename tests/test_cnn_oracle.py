"""Pins for the CNN oracle: its explicit backward equals torch autograd of
the FP32 network, circular padding equals periodic wrap-around, and the
1x1x1 special case equals the pointwise MLP (the reference's
pointwise_apply semantics)."""
import numpy as np
import torch

from oracle import cnn as OC

SHAPES = [(8, 4), (8, 8), (8, 8), (4, 8)]


def params(seed, width=8):
    r = np.random.default_rng(seed)
    h = [r.standard_normal(27 * co * ci).astype(np.float32) * 0.2 for co, ci in SHAPES]
    h += [r.standard_normal(width).astype(np.float32) * 0.1 for _ in range(3)]
    return OC.split_params(np.concatenate(h), SHAPES, width)


def test_backward_matches_autograd():
    n, B = 6, 2
    Ws, bs = params(0)
    x = torch.tensor(np.random.default_rng(1).standard_normal((B, n**3, 4)), dtype=torch.float32)
    # autograd version
    Wt = [OC.torch_weight(W).requires_grad_(True) for W in Ws]
    bt = [torch.tensor(b).requires_grad_(True) for b in bs]
    xs = OC.to_ncdhw(x, n).requires_grad_(True)
    a = xs
    for l in range(3):
        a = OC.gelu(OC.conv(a, Wt[l]) + bt[l].view(1, -1, 1, 1, 1))
    y = OC.to_nvc(OC.conv(a, Wt[3]))
    tgt = torch.randn(B, n**3, 4, generator=torch.Generator().manual_seed(2))
    loss = ((y - tgt) ** 2).mean()
    loss.backward()
    # explicit version
    yy, store = OC.forward(Ws, bs, x, n, emulate=False)
    assert torch.allclose(yy, y.detach(), atol=1e-5)
    _, dy = OC.mse(yy, tgt, 4)
    gW, gb, dx = OC.backward(store, dy, n, emulate=False)
    rel = lambda a, b: float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
    for l in range(4):
        ref = Wt[l].grad.permute(2, 3, 4, 0, 1).reshape(27, *Wt[l].shape[:2])
        assert rel(gW[l], ref) < 1e-5, l
    for l in range(3):
        assert rel(gb[l], bt[l].grad) < 1e-5
    assert rel(dx, OC.to_nvc(xs.grad)) < 1e-5


def test_circular_padding_is_periodic():
    n = 5
    x = torch.randn(1, 1, n, n, n)
    W = torch.zeros(27, 1, 1)
    W[0, 0, 0] = 1.0  # tap (dz,dy,dx) = (-1,-1,-1)
    y = OC.conv(x, OC.torch_weight(W.numpy()))
    assert torch.allclose(y, torch.roll(x, shifts=(1, 1, 1), dims=(2, 3, 4)))


def test_pointwise_special_case():
    n, B = 4, 1
    Ws, bs = params(3)
    for W in Ws:
        W[np.arange(27) != 13] = 0.0
    x = np.random.default_rng(4).standard_normal((B, n**3, 4)).astype(np.float32)
    y, _ = OC.forward(Ws, bs, torch.tensor(x), n, emulate=False)
    ref = OC.pointwise_mlp(Ws, bs, x)
    assert np.allclose(y.numpy(), ref, rtol=1e-4, atol=1e-5)
