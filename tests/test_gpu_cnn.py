"""GPU parity of the BF16-MFMA CNN closure against the torch-CPU oracle that
rounds at the same BF16 storage points.  Tolerances: forward 1e-2 and
gradients 3e-2 relative L2 (BF16 storage of activations and data gradients,
FP32 accumulation in a different order)."""
import numpy as np
import pytest
import torch

from helpers import print_case, rel
from oracle import cnn as OC

pytestmark = pytest.mark.gpu


def make(n=64, batch=1, seed=0):
    import symcode as sc

    return sc.cnn.ConvClosureNet(n, batch_max=batch, cin=4, width=64, cout=6, seed=seed)


def rand_input(net, batch, seed):
    r = np.random.default_rng(seed)
    x = np.zeros((batch, net.nvox, net.cin_p), np.float32)
    x[..., :4] = r.standard_normal((batch, net.nvox, 4))
    return x


def test_cnn_forward_backward_parity():
    print_case("cnn parity")
    net = make()
    n, B = net.n, 1
    h = net.get_params()
    Ws, bs = net.unflatten(h)
    bs = [np.random.default_rng(9 + i).standard_normal(64).astype(np.float32) * 0.05 for i in range(3)]
    h[-192:] = np.concatenate(bs)
    net.set_params(h)
    x = rand_input(net, B, 1)
    xt = torch.tensor(x).to(torch.bfloat16).cuda()
    y = net.forward(xt)
    yo, store = OC.forward(Ws, bs, torch.tensor(x), n)
    assert rel(y.cpu().numpy(), yo.numpy()) < 1e-2
    tgt = np.random.default_rng(2).standard_normal((B, net.nvox, 6)).astype(np.float32) * 0.1
    loss, dy = net.mse(y, torch.tensor(tgt).cuda())
    lo, dyo = OC.mse(yo, tgt, 6)
    assert loss == pytest.approx(lo, rel=1e-2)
    net.zero_grad()
    dx = net.backward(xt, dy, want_dx=True)
    gW, gb, dxo = OC.backward(store, dyo, n)
    g = net.get_grads()
    GW, GB = net.unflatten(g)
    for l in range(4):
        assert rel(GW[l], gW[l].numpy()) < 3e-2, l
    for l in range(3):
        assert rel(GB[l], gb[l].numpy()) < 3e-2, l
    assert rel(dx.cpu().numpy(), dxo.numpy()) < 3e-2


def test_cnn_pointwise_special_case():
    """only the centre tap non-zero: the network is the reference's pointwise
    MLP (pointwise_apply + GELU, src/nets.jl:63-81)"""
    print_case("cnn pointwise")
    net = make()
    h = net.get_params()
    Ws, bs = net.unflatten(h)
    for W in Ws:
        W[np.arange(27) != 13] = 0.0
    net.set_params(h)
    x = rand_input(net, 1, 3)
    xb = torch.tensor(x).to(torch.bfloat16)
    y = net.forward(xb.cuda()).cpu().numpy()
    Wr = [torch.tensor(W).to(torch.bfloat16).float().numpy() for W in Ws]
    ref = OC.pointwise_mlp(Wr, bs, xb.float().numpy())
    assert rel(y, ref) < 1e-2


def test_cnn_training_reduces_loss():
    print_case("cnn train")
    net = make(seed=4)
    x = torch.tensor(rand_input(net, 1, 5)).to(torch.bfloat16).cuda()
    # a smooth target: a fixed random conv net's output
    teacher = make(seed=77)
    tgt = teacher.forward(x)[..., :6].contiguous()
    losses = [net.train_step(x, tgt, lr=1e-3) for _ in range(30)]
    print("losses", losses[0], losses[9], losses[19], losses[-1])
    assert losses[-1] < 0.7 * losses[0], (losses[0], losses[-1])


def test_cnn_pack_input():
    import symcode as sc

    print_case("cnn pack")
    net = make()
    n = net.n
    f = torch.randn(3, n, n, n, dtype=torch.float64, device="cuda")
    x = net.pack_input(f, scale=2.0, extra=0.75)
    xc = x.float().cpu().numpy()[0]
    ref = (2.0 * f).reshape(3, -1).T.cpu().numpy()
    assert np.allclose(xc[:, :3], torch.tensor(ref).to(torch.bfloat16).float().numpy())
    assert np.all(xc[:, 3] == np.float32(torch.tensor(0.75).to(torch.bfloat16).float()))
    assert np.all(xc[:, 4:] == 0)


def test_cnn_parity_n128():
    """two 64-voxel x segments per row: forward and gradients at n = 128"""
    print_case("cnn parity n128")
    net = make(n=128, seed=5)
    n = net.n
    h = net.get_params()
    Ws, bs = net.unflatten(h)
    x = rand_input(net, 1, 6)
    xt = torch.tensor(x).to(torch.bfloat16).cuda()
    y = net.forward(xt)
    yo, store = OC.forward(Ws, bs, torch.tensor(x), n)
    assert rel(y.cpu().numpy(), yo.numpy()) < 1e-2
    tgt = np.random.default_rng(7).standard_normal((1, net.nvox, 6)).astype(np.float32) * 0.1
    _, dy = net.mse(y, torch.tensor(tgt).cuda())
    _, dyo = OC.mse(yo, tgt, 6)
    net.zero_grad()
    dx = net.backward(xt, dy, want_dx=True)
    gW, gb, dxo = OC.backward(store, dyo, n)
    GW, GB = net.unflatten(net.get_grads())
    for l in range(4):
        assert rel(GW[l], gW[l].numpy()) < 3e-2, l
    for l in range(3):
        assert rel(GB[l], gb[l].numpy()) < 3e-2, l
    assert rel(dx.cpu().numpy(), dxo.numpy()) < 3e-2


@pytest.mark.parametrize("cin", [3, 7])
def test_cnn_parity_wide_output(cin):
    """12 output channels (16-row output tile, the weight gradient without
    plane packing) and 3 or 7 input channels (the input layer's weight
    gradient with 4 taps x 4 channels per MFMA tile, or 2 taps x 8):
    forward and gradients"""
    import symcode as sc

    print_case(f"cnn parity cout12 cin{cin}")
    net = sc.cnn.ConvClosureNet(64, batch_max=1, cin=cin, width=64, cout=12, seed=13)
    n = net.n
    assert (net.cin_p, net.cout_p) == (8, 16)
    Ws, bs = net.unflatten(net.get_params())
    r = np.random.default_rng(14)
    x = np.zeros((1, net.nvox, net.cin_p), np.float32)
    x[..., :cin] = r.standard_normal((1, net.nvox, cin))
    xt = torch.tensor(x).to(torch.bfloat16).cuda()
    y = net.forward(xt)
    yo, store = OC.forward(Ws, bs, torch.tensor(x), n)
    assert rel(y.cpu().numpy(), yo.numpy()) < 1e-2
    tgt = r.standard_normal((1, net.nvox, 12)).astype(np.float32) * 0.1
    _, dy = net.mse(y, torch.tensor(tgt).cuda())
    _, dyo = OC.mse(yo, tgt, 12)
    net.zero_grad()
    dx = net.backward(xt, dy, want_dx=True)
    gW, gb, dxo = OC.backward(store, dyo, n)
    GW, GB = net.unflatten(net.get_grads())
    for l in range(4):
        assert rel(GW[l], gW[l].numpy()) < 3e-2, l
    for l in range(3):
        assert rel(GB[l], gb[l].numpy()) < 3e-2, l
    assert rel(dx.cpu().numpy(), dxo.numpy()) < 3e-2


def test_cnn_tracefree_parity():
    """trace removal on the output (the reference nets' last layer): zero
    trace, forward and gradients against the oracle's projected network"""
    import symcode as sc

    print_case("cnn tracefree")
    net = sc.cnn.ConvClosureNet(64, batch_max=1, cin=4, width=64, cout=6, seed=8, tracefree=True)
    n = net.n
    Ws, bs = net.unflatten(net.get_params())
    x = rand_input(net, 1, 11)
    xt = torch.tensor(x).to(torch.bfloat16).cuda()
    y = net.forward(xt)
    yc = y.cpu().numpy()
    assert np.abs(yc[..., :3].sum(-1)).max() < 1e-5 * np.abs(yc[..., :3]).max()
    yo, store = OC.forward(Ws, bs, torch.tensor(x), n, tracefree_out=True)
    assert rel(yc, yo.numpy()) < 1e-2
    # a target with a trace: the projection must reach the gradient
    tgt = np.random.default_rng(12).standard_normal((1, net.nvox, 6)).astype(np.float32) * 0.1
    _, dy = net.mse(y, torch.tensor(tgt).cuda())
    _, dyo = OC.mse(yo, tgt, 6)
    net.zero_grad()
    dx = net.backward(xt, dy, want_dx=True)
    gW, gb, dxo = OC.backward(store, dyo, n)
    GW, GB = net.unflatten(net.get_grads())
    for l in range(4):
        assert rel(GW[l], gW[l].numpy()) < 3e-2, l
    for l in range(3):
        assert rel(GB[l], gb[l].numpy()) < 3e-2, l
    assert rel(dx.cpu().numpy(), dxo.numpy()) < 3e-2
    with pytest.raises(ValueError):
        sc.cnn.ConvClosureNet(64, batch_max=1, cin=4, width=64, cout=5, tracefree=True)


def test_cnn_relative_loss():
    """the reference's relative loss and its gradient on the GPU output"""
    print_case("cnn relative loss")
    net = make(seed=13, batch=2)
    x = torch.tensor(rand_input(net, 2, 14)).to(torch.bfloat16).cuda()
    y = net.forward(x)
    tgt = np.random.default_rng(15).standard_normal((2, net.nvox, 6)).astype(np.float32) * 0.3
    loss, dy = net.loss(y, torch.tensor(tgt).cuda(), "relative")
    lo, dyo = OC.rel_loss(y.cpu(), tgt, 6)
    assert loss == pytest.approx(lo, rel=1e-9)
    assert rel(dy.float().cpu().numpy(), dyo.numpy()) < 5e-3
    with pytest.raises(ValueError):
        net.loss(y, torch.tensor(tgt).cuda(), "l1")


def test_cnn_adamw_clip_matches_oracle():
    """three ClipNorm + AdamW steps (with decay) on the network's own
    gradients equal the oracle's optimiser applied to the same gradients"""
    from oracle import optim as O

    print_case("cnn adamw clip")
    net = make(seed=16)
    x = torch.tensor(rand_input(net, 1, 17)).to(torch.bfloat16).cuda()
    shapes = net.layer_shapes()
    bounds = [0]
    for co, ci in shapes:
        bounds.append(bounds[-1] + 27 * co * ci)
    for _ in range(3):
        bounds.append(bounds[-1] + net.width)
    st = O.AdamWState(net.nparams)
    w = net.get_params().astype(np.float64)
    for k, (clip, wd) in enumerate([(1.0, 0.0), (0.05, 0.01), (1e3, 0.0)]):
        net.zero_grad()
        y = net.forward(x)
        tgt = torch.full((1, net.nvox, 6), 0.1 * (k + 1), device="cuda")
        _, dy = net.loss(y, tgt, "relative")
        net.backward(x, dy)
        gr = net.get_grads().astype(np.float64)
        net.adamw(1e-3, weight_decay=wd, clip=clip)
        w = O.clipnorm_adamw(w, gr, st, bounds, 1e-3, weight_decay=wd, clip=clip)
        got = net.get_params().astype(np.float64)
        assert np.abs(got - w).max() < 1e-6, (k, np.abs(got - w).max())
        assert np.array_equal(net.get_grads(), gr.astype(np.float32))  # gradient kept
