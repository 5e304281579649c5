"""GPU parity of the pointwise MLP closure (sc_pmlp_*, closure kind 7) against
oracle/pmlp.py: the training pair's features and the FP32 network (forward,
relative and MSE losses, every parameter gradient, ClipNorm + AdamW) within
FP32 tolerances; fullchain's FP64 inference and les! with it within 1e-12
relative on FP64 fields (1e-5 on FP32 fields)."""
import math

import numpy as np
import pytest
import torch

from helpers import oracle_solenoidal, print_case, rand_phys, rel
from oracle import closures as OC
from oracle import optim as OO
from oracle import pmlp as OP
from oracle import spectral as S

pytestmark = pytest.mark.gpu


def net_of(D, widths=(7, 9, 5), use_redelta=False, seed=1, nvox_max=4096):
    from symcode.pmlp import PointwiseMLP

    return PointwiseMLP(D, widths, use_redelta=use_redelta, seed=seed, nvox_max=nvox_max)


def oracle_params(net):
    return OP.unflat(net.get_params().astype(np.float64), net.layer_shapes())


def test_create_info_and_init():
    print_case("pmlp create", 3)
    net = net_of(3, (22, 36, 44), use_redelta=True, seed=5)
    assert (net.D, net.nin, net.widths, net.nout) == (3, 10, (22, 36, 44), 6)
    shapes = net.layer_shapes()
    assert net.nparams == OP.array_bounds(shapes)[-1] == 10 * 22 + 22 + 22 * 36 + 36 + 36 * 44 + 44 + 44 * 6
    p = oracle_params(net)
    for l, (ci, co) in enumerate(shapes):
        W = p[f"W{l + 1}"]
        lim = math.sqrt(6 / (ci + co))
        assert np.abs(W).max() <= lim and np.abs(W).max() > 0.8 * lim
        assert abs(W.mean()) < 0.2 * lim
        if l < 3:
            assert np.all(p[f"b{l + 1}"] == 0)
    h = net.get_params()
    h2 = np.random.default_rng(0).standard_normal(h.size).astype(np.float32)
    net.set_params(h2)
    assert np.array_equal(net.get_params(), h2)
    assert np.array_equal(net_of(3, (22, 36, 44), True, seed=5).get_params(), h)
    # same_as_equi multiplies the widths by |G|
    from symcode.pmlp import PointwiseMLP

    assert PointwiseMLP(2, (1, 2, 3), same_as_equi=True, nvox_max=64).widths == (8, 16, 24)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_features_parity(dtype):
    import symcode as sc

    print_case("pmlp features", dtype)
    go = S.Grid(3, 16)
    g = sc.Grid(3, 16, dtype=dtype)
    net = net_of(3, use_redelta=True)
    u = oracle_solenoidal(go, 3, trunc=True)
    GG = sc.gradient_phys(g.upload_spectral(u), g)
    tau = g.upload_physical(np.stack([rand_phys(go.phys_shape, 20 + i) for i in range(6)]))
    delta, z = 0.35, -0.6
    x, y = net.features(GG, delta, tau, z)
    eps = np.finfo(np.float64 if dtype == torch.float64 else np.float32).eps
    xo, yo = OP.features(g.download_physical(GG), g.download_physical(tau), delta, z, eps=eps)
    assert np.allclose(x.cpu().numpy(), xo.astype(np.float32), rtol=3e-7, atol=1e-30)
    assert np.allclose(y.cpu().numpy(), yo.astype(np.float32), rtol=3e-7, atol=1e-30)


@pytest.mark.parametrize("D,use_redelta", [(3, False), (3, True), (2, False)])
def test_forward_loss_backward_parity(D, use_redelta):
    print_case("pmlp forward backward", D, use_redelta)
    net = net_of(D, use_redelta=use_redelta)
    nv = 3000
    rng = np.random.default_rng(D)
    x = rng.standard_normal((nv, net.nin)).astype(np.float32)
    t = rng.standard_normal((nv, net.nout)).astype(np.float32)
    xd, td = torch.tensor(x, device="cuda"), torch.tensor(t, device="cuda")
    y = net.forward(xd)
    p = oracle_params(net)
    yo, cache = OP.forward(p, x.astype(np.float64), D)
    assert rel(y.cpu().numpy(), yo) < 2e-6
    k = 3 if D == 3 else 2
    assert np.abs(y.cpu().numpy()[:, :k].sum(axis=1)).max() < 1e-5
    for kind, fn in (("mse", OP.mse_loss), ("relative", OP.relative_loss)):
        net.zero_grad()
        y = net.forward(xd)
        loss, dy = net.loss(y, td, kind)
        lo, dyo = fn(yo, t.astype(np.float64))
        assert loss == pytest.approx(lo, rel=1e-5)
        assert rel(dy.cpu().numpy(), dyo) < 1e-5
        net.backward(xd, dy)
        go_ = OP.backward(p, cache, dyo, D)
        gflat = net.get_grads()
        gg = OP.unflat(gflat, net.layer_shapes())
        for name in go_:
            assert rel(gg[name], go_[name]) < 2e-5, (kind, name)
    # gradients accumulate until zero_grad
    net.zero_grad()
    for _ in range(2):
        y = net.forward(xd)
        _, dy = net.loss(y, td, "relative")
        net.backward(xd, dy)
    assert rel(net.get_grads(), 2 * gflat) < 1e-6


def test_adamw_steps_parity():
    print_case("pmlp adamw")
    net = net_of(3, seed=2)
    rng = np.random.default_rng(9)
    x = torch.tensor(rng.standard_normal((2000, 9)).astype(np.float32), device="cuda")
    t = torch.tensor(rng.standard_normal((2000, 6)).astype(np.float32), device="cuda")
    shapes = net.layer_shapes()
    w = net.get_params().astype(np.float64)
    st = OO.AdamWState(w.size)
    for lr, clip in ((1e-2, 0.05), (5e-3, 0.05), (1e-3, 0.0)):
        net.set_params(w.astype(np.float32))
        w32 = net.get_params().astype(np.float64)
        net.zero_grad()
        y = net.forward(x)
        _, dy = net.loss(y, t, "relative")
        net.backward(x, dy)
        g = net.get_grads().astype(np.float64)
        net.adamw(lr, clip=clip, weight_decay=1e-2)
        w = OO.clipnorm_adamw(w32, g, st, OP.array_bounds(shapes), lr, weight_decay=1e-2, clip=clip)
        assert np.allclose(net.get_params(), w, rtol=1e-5, atol=1e-7)


# widths up to 64 run the fused MFMA network (one, three and four tiles of 16
# channels; 30^3 = 27000 rows ends in a partial row tile), wider ones one GEMM
# per layer
@pytest.mark.parametrize("D,n,use_redelta,widths", [(3, 16, False, (11, 13, 7)), (3, 16, True, (11, 13, 7)),
                                                    (2, 32, True, (11, 13, 7)), (3, 30, True, (22, 36, 44)),
                                                    (2, 30, False, (22, 36, 44)), (3, 16, True, (44, 64, 64)),
                                                    (3, 16, True, (70, 13, 7))])
def test_closure_parity_fp64(D, n, use_redelta, widths):
    import symcode as sc
    from symcode.pmlp import PointwiseClosure

    print_case("pmlp closure fp64", D, n, use_redelta, widths)
    go = S.Grid(D, n)
    g = sc.Grid(D, n, dtype=torch.float64)
    net = net_of(D, widths, use_redelta=use_redelta, seed=4)
    delta, visc, norm = 2 * np.pi / n * 2, 2e-3, (1.1, 0.8)
    model = PointwiseClosure(net, g, delta, visc, norm if use_redelta else None)
    u = oracle_solenoidal(go, 31, trunc=True)
    ud = g.upload_spectral(u)
    tau = g.download_physical(model(ud))
    p = oracle_params(net)
    GGo = OC.gradient_phys(u, go)
    to = OP.fullchain(p, GGo, delta, D, redelta=(visc, *norm) if use_redelta else None)
    assert rel(tau, to) < 1e-12
    # les! with closure 7
    du = sc.les_rhs(g.vectorfield(), ud, g, model, visc)
    duo = OC.les_rhs(u, go, visc, lambda u_, GG_: OP.fullchain(p, GG_, delta, D, redelta=(visc, *norm)
                                                              if use_redelta else None))
    assert rel(g.download_spectral(du), duo) < 1e-12


def test_closure_parity_fp32():
    import symcode as sc
    from symcode.pmlp import PointwiseClosure

    print_case("pmlp closure fp32", 3, 16)
    go = S.Grid(3, 16)
    g = sc.Grid(3, 16, dtype=torch.float32)
    net = net_of(3, (11, 13, 7), use_redelta=True, seed=6)
    delta, visc, norm = 2 * np.pi / 16 * 2, 2e-3, (1.1, 0.8)
    model = PointwiseClosure(net, g, delta, visc, norm)
    u = oracle_solenoidal(go, 33, trunc=True)
    tau = g.download_physical(model(g.upload_spectral(u)))
    to = OP.fullchain(oracle_params(net), OC.gradient_phys(u, go), delta, 3, eps=np.finfo(np.float32).eps,
                      redelta=(visc, *norm))
    assert rel(tau, to) < 1e-5


def test_closure_chunks_and_errors():
    """a 128^3 field runs in two 2^20-voxel chunks; the adjoint rejects kind 7"""
    import symcode as sc
    from symcode import SymcodeError
    from symcode.pmlp import PointwiseClosure

    print_case("pmlp closure 128^3 chunks", 3, 128)
    go = S.Grid(3, 128)
    g = sc.Grid(3, 128, dtype=torch.float64)
    net = net_of(3, (6, 10, 14), seed=8, nvox_max=256)
    delta = 2 * np.pi / 128 * 2
    model = PointwiseClosure(net, g, delta)
    u = oracle_solenoidal(go, 35, trunc=True)
    ud = g.upload_spectral(u)
    tau = g.download_physical(model(ud))
    to = OP.fullchain(oracle_params(net), OC.gradient_phys(u, go), delta, 3)
    assert rel(tau, to) < 1e-12
    with pytest.raises(SymcodeError):
        sc.les_rhs_vjp(g.vectorfield(), ud, g.vectorfield(), g, model, 1e-3)
    with pytest.raises(SymcodeError):
        net.forward(torch.zeros((257, 9), dtype=torch.float32, device="cuda"))
    with pytest.raises(ValueError):
        PointwiseClosure(net_of(3, use_redelta=True), g, delta)


def test_training_on_filtered_pairs():
    """create_dataloader's pairs from filtered 32^3 fields, trained with the
    reference's semantics; the first step equals the oracle's and the loss
    falls (by 10x on a teacher network's targets); the trained closure then
    runs in a short LES rollout"""
    import symcode as sc
    from symcode import pmlp as PM
    from symcode import training as T
    from symcode.training import Pair

    print_case("pmlp training 32->16")
    god = S.Grid(3, 32)
    gd = sc.Grid(3, 32, dtype=torch.float64)
    gl = sc.Grid(3, 16, dtype=torch.float64)
    delta, visc = 2 * np.pi / 16 * 2, 5e-3
    pairs = []
    for seed in (71, 72, 73):
        u = oracle_solenoidal(god, seed, trunc=True)
        ub, tau = sc.sfs(gd.upload_spectral(u), gd, gl, delta)
        pairs.append(Pair(ub, tau, delta, T.filter_reynolds(ub, gl, visc, delta)))
    norm = PM.redelta_norm(pairs)
    logre = np.log([p.redelta for p in pairs])
    assert norm == pytest.approx((logre.mean(), logre.std(ddof=1)), rel=1e-12)
    net = PM.PointwiseMLP(3, (10, 18, 32), use_redelta=True, seed=0, nvox_max=3 * 16**3)
    X, Y = PM.pointwise_pairs(net, pairs, gl, norm)
    assert X.shape == (3 * 16**3, 10) and Y.shape == (3 * 16**3, 6)
    # the pair of the second snapshot against the oracle's dataloader
    gol = S.Grid(3, 16)
    ubo, tauo = gl.download_spectral(pairs[1].ubar), gl.download_spectral(pairs[1].tau)
    GGo = OC.gradient_phys(ubo, gol)
    tph = np.stack([S.to_phys(S.twothirds(t, gol), gol) for t in tauo])
    z = (math.log(pairs[1].redelta) - norm[0]) / norm[1]
    xo, yo = OP.features(GGo, tph, delta, z)
    assert rel(X[16**3:2 * 16**3].cpu().numpy(), xo) < 1e-6
    assert rel(Y[16**3:2 * 16**3].cpu().numpy(), yo) < 1e-6
    # one step against the oracle
    p = oracle_params(net)
    w0 = net.get_params().astype(np.float64)
    x64, y64 = X.cpu().numpy().astype(np.float64), Y.cpu().numpy().astype(np.float64)
    yh, cache = OP.forward(p, x64, 3)
    lo, dyo = OP.relative_loss(yh, y64)
    go_ = OP.backward(p, cache, dyo, 3)
    l0 = net.train_step(X, Y, 1e-3, loss="relative", clip=1.0)
    assert l0 == pytest.approx(lo, rel=1e-5)
    w1 = OO.clipnorm_adamw(w0, OP.flat(go_, net.layer_shapes()), OO.AdamWState(w0.size),
                           OP.array_bounds(net.layer_shapes()), 1e-3, clip=1.0)
    assert np.allclose(net.get_params(), w1, rtol=1e-4, atol=1e-6)
    # pointwise, the stress of these filtered random fields is barely
    # predictable from the local gradient: training only has to lower it
    losses = T.reference_train(net, X, Y, 150, lr=3e-3)
    print("loss", l0, losses[0], losses[-1], flush=True)
    assert losses[-1] < min(losses[0], l0)
    # a teacher network's outputs on the same inputs are learnable exactly
    teacher = PM.PointwiseMLP(3, (10, 18, 32), use_redelta=True, seed=9, nvox_max=3 * 16**3)
    Yt = teacher.forward(X).clone()
    student = PM.PointwiseMLP(3, (10, 18, 32), use_redelta=True, seed=1, nvox_max=3 * 16**3)
    lt = T.reference_train(student, X, Yt, 150, lr=3e-3)
    print("teacher-student loss", lt[0], lt[-1], flush=True)
    assert lt[-1] < 0.1 * lt[0]
    model = PM.PointwiseClosure(net, gl, delta, visc, norm)
    u0 = pairs[0].ubar.clone()
    nred = sc.solve_les(u0, [0.0, 0.02], gl, visc, model, 0.3, forced=False)
    assert nred == 2 and torch.isfinite(u0).all()


# ------------------------------------------------------------------ TBNN
def tbnn_of(D, widths=(7, 9, 5), use_redelta=False, seed=1, nvox_max=4096):
    from symcode.pmlp import TBNN

    return TBNN(D, widths, use_redelta=use_redelta, seed=seed, nvox_max=nvox_max)


@pytest.mark.parametrize("D,n,dtype", [(3, 16, torch.float64), (3, 16, torch.float32), (2, 32, torch.float64)])
def test_tbnn_features_parity(D, n, dtype):
    import symcode as sc

    print_case("tbnn features", D, n, dtype)
    go = S.Grid(D, n)
    g = sc.Grid(D, n, dtype=dtype)
    net = tbnn_of(D, use_redelta=True, nvox_max=n**D)
    assert net.row == (5 if D == 3 else 2) + 1 + (42 if D == 3 else 6)
    u = oracle_solenoidal(go, 7, trunc=True)
    GG = sc.gradient_phys(g.upload_spectral(u), g)
    tau = g.upload_physical(np.stack([rand_phys(go.phys_shape, 30 + i) for i in range(go.tensordim)]))
    delta, z = 0.35, 0.4
    x, y = net.features(GG, delta, tau, z)
    eps = np.finfo(np.float64 if dtype == torch.float64 else np.float32).eps
    GGo = g.download_physical(GG)
    xo = OP.tb_rows(GGo, D, z, eps)
    _, yo = OP.features(GGo, g.download_physical(tau), delta, None, eps)
    assert rel(x.cpu().numpy(), xo) < 1e-6
    assert rel(y.cpu().numpy(), yo) < 1e-6


@pytest.mark.parametrize("D,use_redelta", [(3, True), (2, False)])
def test_tbnn_forward_loss_backward_parity(D, use_redelta):
    print_case("tbnn forward backward", D, use_redelta)
    net = tbnn_of(D, use_redelta=use_redelta)
    rng = np.random.default_rng(D + 10)
    GG = rng.standard_normal((D * D, 3000))
    x = OP.tb_rows(GG, D, 0.2 if use_redelta else None).astype(np.float32)
    t = rng.standard_normal((3000, net.nout)).astype(np.float32)
    xd, td = torch.tensor(x, device="cuda"), torch.tensor(t, device="cuda")
    p = oracle_params(net)
    net.zero_grad()
    y = net.forward(xd)
    mo, cache = OP.tbnn_forward(p, x.astype(np.float64), D, net.nin)
    assert rel(y.cpu().numpy(), mo) < 2e-6
    loss, dy = net.loss(y, td, "relative")
    lo, dmo = OP.relative_loss(mo, t.astype(np.float64))
    assert loss == pytest.approx(lo, rel=1e-5)
    net.backward(xd, dy)
    go_ = OP.tbnn_backward(p, cache, dmo)
    gg = OP.unflat(net.get_grads(), net.layer_shapes())
    for name in go_:
        assert rel(gg[name], go_[name]) < 2e-5, name


# up to three tiles of 16 channels: one fused kernel (30^3 ends in a partial
# row tile); 64 channels: the chunked path
@pytest.mark.parametrize("D,n,use_redelta,widths", [(3, 16, True, (11, 13, 7)), (2, 32, False, (11, 13, 7)),
                                                    (3, 30, True, (22, 36, 44)), (2, 30, True, (22, 36, 44)),
                                                    (3, 16, False, (46, 64, 64))])
def test_tbnn_closure_parity_fp64(D, n, use_redelta, widths):
    import symcode as sc
    from symcode.pmlp import PointwiseClosure

    print_case("tbnn closure fp64", D, n, use_redelta, widths)
    go = S.Grid(D, n)
    g = sc.Grid(D, n, dtype=torch.float64)
    net = tbnn_of(D, widths, use_redelta=use_redelta, seed=4)
    delta, visc, norm = 2 * np.pi / n * 2, 2e-3, (1.1, 0.8)
    model = PointwiseClosure(net, g, delta, visc, norm if use_redelta else None)
    u = oracle_solenoidal(go, 37, trunc=True)
    ud = g.upload_spectral(u)
    p = oracle_params(net)
    re = (visc, *norm) if use_redelta else None
    to = OP.tbnn_model(p, OC.gradient_phys(u, go), delta, D, redelta=re)
    assert rel(g.download_physical(model(ud)), to) < 1e-12
    du = sc.les_rhs(g.vectorfield(), ud, g, model, visc)
    duo = OC.les_rhs(u, go, visc, lambda u_, GG_: OP.tbnn_model(p, GG_, delta, D, redelta=re))
    assert rel(g.download_spectral(du), duo) < 1e-12
    with pytest.raises(ValueError):
        PointwiseClosure(net, g, delta, visc, norm, symmetrize=True)


def test_tbnn_closure_parity_fp32():
    import symcode as sc
    from symcode.pmlp import PointwiseClosure

    print_case("tbnn closure fp32", 3, 16)
    go = S.Grid(3, 16)
    g = sc.Grid(3, 16, dtype=torch.float32)
    net = tbnn_of(3, (11, 13, 7), use_redelta=True, seed=6)
    delta, visc, norm = 2 * np.pi / 16 * 2, 2e-3, (1.1, 0.8)
    model = PointwiseClosure(net, g, delta, visc, norm)
    u = oracle_solenoidal(go, 41, trunc=True)
    tau = g.download_physical(model(g.upload_spectral(u)))
    to = OP.tbnn_model(oracle_params(net), OC.gradient_phys(u, go), delta, 3, eps=np.finfo(np.float32).eps,
                       redelta=(visc, *norm))
    assert rel(tau, to) < 1e-5


def test_tbnn_training_teacher():
    """a teacher TBNN's outputs on the rows of a gradient field are learnable:
    the reference semantics lower the relative loss tenfold"""
    import symcode as sc
    from symcode import training as T

    print_case("tbnn teacher training")
    go = S.Grid(3, 16)
    g = sc.Grid(3, 16, dtype=torch.float64)
    GG = sc.gradient_phys(g.upload_spectral(oracle_solenoidal(go, 8, trunc=True)), g)
    teacher = tbnn_of(3, (10, 18, 32), seed=9, nvox_max=16**3)
    X = teacher.features(GG, 0.3)
    Yt = teacher.forward(X).clone()
    student = tbnn_of(3, (10, 18, 32), seed=2, nvox_max=16**3)
    lt = T.reference_train(student, X, Yt, 150, lr=3e-3)
    print("teacher-student loss", lt[0], lt[-1], flush=True)
    assert lt[-1] < 0.1 * lt[0]


# ------------------------------------------- symmetrize_pointwise (convsym)
@pytest.mark.parametrize("D,n,use_redelta", [(3, 16, True), (2, 32, False)])
def test_symmetrized_closure_parity_fp64(D, n, use_redelta):
    import symcode as sc
    from symcode.pmlp import PointwiseClosure, symmetrize_pointwise

    print_case("symmetrized mlp closure fp64", D, n, use_redelta)
    go = S.Grid(D, n)
    g = sc.Grid(D, n, dtype=torch.float64)
    net = net_of(D, (11, 13, 7), use_redelta=use_redelta, seed=4)
    delta, visc, norm = 2 * np.pi / n * 2, 2e-3, (1.1, 0.8)
    plain = PointwiseClosure(net, g, delta, visc, norm if use_redelta else None)
    model = symmetrize_pointwise(plain)
    u = oracle_solenoidal(go, 39, trunc=True)
    ud = g.upload_spectral(u)
    p = oracle_params(net)
    re = (visc, *norm) if use_redelta else None
    GGo = OC.gradient_phys(u, go)
    fc = lambda G: OP.fullchain(p, G, delta, D, redelta=re)
    to = OP.symmetrize(fc, GGo, D)
    tau = g.download_physical(model(ud))
    assert rel(tau, to) < 1e-12
    assert rel(tau, fc(GGo)) > 1e-3  # the average differs from the plain network
    du = sc.les_rhs(g.vectorfield(), ud, g, model, visc)
    duo = OC.les_rhs(u, go, visc, lambda u_, GG_: OP.symmetrize(fc, GG_, D))
    assert rel(g.download_spectral(du), duo) < 1e-12
    # attaching the plain closure switches the average off again
    assert rel(g.download_physical(plain(ud)), fc(GGo)) < 1e-12


def test_symmetrized_closure_parity_fp32():
    import symcode as sc
    from symcode.pmlp import PointwiseClosure, symmetrize_pointwise

    print_case("symmetrized mlp closure fp32", 3, 16)
    go = S.Grid(3, 16)
    g = sc.Grid(3, 16, dtype=torch.float32)
    net = net_of(3, (22, 36, 44), use_redelta=True, seed=5)
    delta, visc, norm = 2 * np.pi / 16 * 2, 2e-3, (1.1, 0.8)
    model = symmetrize_pointwise(PointwiseClosure(net, g, delta, visc, norm))
    u = oracle_solenoidal(go, 43, trunc=True)
    tau = g.download_physical(model(g.upload_spectral(u)))
    p = oracle_params(net)
    fc = lambda G: OP.fullchain(p, G, delta, 3, eps=np.finfo(np.float32).eps, redelta=(visc, *norm))
    assert rel(tau, OP.symmetrize(fc, OC.gradient_phys(u, go), 3)) < 1e-5


# ------------------------------------------------------ equivariant_net
def equi_of(D, widths=(2, 3, 2), use_redelta=False, seed=1, nvox_max=4096):
    from symcode.pmlp import EquivariantNet

    return EquivariantNet(D, widths, use_redelta=use_redelta, seed=seed, nvox_max=nvox_max)


def flat_f(t):
    """column-major flatten of a torch tensor"""
    return t.permute(*reversed(range(t.dim()))).reshape(-1)


def test_equi_create_and_init():
    print_case("equi create", 3)
    net = equi_of(3, (4, 5, 8), use_redelta=True, seed=3)
    sh = net.compact_shapes()
    assert net.nreg == 48 and net.nparams == sum(int(np.prod(s)) for _, s in sh)
    assert net.nnet == 10 * 192 + 192 + 192 * 240 + 240 + 240 * 384 + 384 + 384 * 9
    h = net.get_params()
    o = 0
    for name, shp in sh:
        a = h[o:o + int(np.prod(shp))]
        o += a.size
        if name.endswith("bias"):
            assert np.all(a == 0)
        else:
            fan = {"lift.weight": 9 + 4, "lift.weight_re": 1 + 4, "mid_1.weight": 48 * 9,
                   "mid_2.weight": 48 * 13, "sink.weight": 9 + 8}[name]
            assert np.abs(a).max() <= math.sqrt(6 / fan) and np.abs(a).max() > 0.7 * math.sqrt(6 / fan)


@pytest.mark.parametrize("D,use_redelta", [(3, True), (2, False)])
def test_equi_forward_loss_backward_parity(D, use_redelta):
    print_case("equi forward backward", D, use_redelta)
    widths = (2, 3, 2)
    net = equi_of(D, widths, use_redelta=use_redelta, seed=2)
    rng = np.random.default_rng(D + 20)
    GG = rng.standard_normal((D * D, 2000))
    x, _ = OP.features(GG, None, 0.3, 0.2 if use_redelta else None)
    x = x.astype(np.float32)
    t = rng.standard_normal((2000, net.nout)).astype(np.float32)
    xd, td = torch.tensor(x, device="cuda"), torch.tensor(t, device="cuda")
    h = net.get_params()
    ps = OP.equi_unflat(h, D, widths, use_redelta)
    for v in ps.values():
        v.requires_grad_(True)
    layers = OP.equi_project(ps, D, OP.weight_synthesis(D))
    yo = OP.equi_net(layers, torch.tensor(x.astype(np.float64)), D)
    tt = torch.tensor(t.astype(np.float64))
    lo = ((yo - tt) ** 2).sum() / ((tt**2).sum() + float(np.finfo(np.float32).eps))
    lo.backward()
    go = torch.cat([flat_f(v.grad) for v in ps.values()]).numpy()
    net.zero_grad()
    y = net.forward(xd)
    assert rel(y.cpu().numpy(), yo.detach().numpy()) < 2e-6
    loss, dy = net.loss(y, td, "relative")
    assert loss == pytest.approx(lo.item(), rel=1e-5)
    net.backward(xd, dy)
    g = net.get_grads()
    # per array, against the whole gradient's scale: a bias gradient sums
    # |G| n_voxel terms that cancel to a small value, and its FP32 rounding is
    # relative to the terms
    o = 0
    for name, shp in net.compact_shapes():
        n = int(np.prod(shp))
        assert np.linalg.norm(g[o:o + n] - go[o:o + n]) < 1e-4 * np.linalg.norm(go), name
        o += n
    assert rel(g, go) < 1e-4
    # one clipped AdamW step over the compact arrays (eight with weight_re)
    w0 = h.astype(np.float64)
    bounds = [0]
    for _, shp in net.compact_shapes():
        bounds.append(bounds[-1] + int(np.prod(shp)))
    w1 = OO.clipnorm_adamw(w0, g.astype(np.float64), OO.AdamWState(w0.size), bounds, 1e-2, weight_decay=1e-2,
                           clip=0.05)
    net.adamw(1e-2, weight_decay=1e-2, clip=0.05)
    assert np.allclose(net.get_params(), w1, rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("D,n,use_redelta", [(3, 16, True), (2, 32, False)])
def test_equi_closure_parity_fp64(D, n, use_redelta):
    import symcode as sc
    from symcode.pmlp import PointwiseClosure

    print_case("equi closure fp64", D, n, use_redelta)
    go = S.Grid(D, n)
    g = sc.Grid(D, n, dtype=torch.float64)
    widths = (2, 3, 2)
    net = equi_of(D, widths, use_redelta=use_redelta, seed=4)
    delta, visc, norm = 2 * np.pi / n * 2, 2e-3, (1.1, 0.8)
    model = PointwiseClosure(net, g, delta, visc, norm if use_redelta else None)
    u = oracle_solenoidal(go, 41, trunc=True)
    ud = g.upload_spectral(u)
    h = net.get_params()
    re = (visc, *norm) if use_redelta else None
    fc = lambda G: OP.equi_fullchain(h, D, widths, G, delta, redelta=re)
    to = fc(OC.gradient_phys(u, go))
    assert rel(g.download_physical(model(ud)), to) < 1e-12
    du = sc.les_rhs(g.vectorfield(), ud, g, model, visc)
    duo = OC.les_rhs(u, go, visc, lambda u_, GG_: fc(GG_))
    assert rel(g.download_spectral(du), duo) < 1e-12
    with pytest.raises(ValueError):
        PointwiseClosure(net, g, delta, visc, norm, symmetrize=True)
