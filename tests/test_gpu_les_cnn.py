"""GPU parity of les! with the learned closure (closure kind 4) and of its
transposes, against oracle/les_cnn.py with the network's BF16 storage points
emulated.  The closure's share is compared at BF16 tolerance: the norm of
(GPU - oracle) against the norm of the closure's contribution (oracle with
closure - oracle without).  The closure-free paths keep the FP64 bar (1e-12)."""
import numpy as np
import pytest
import torch

from helpers import oracle_solenoidal, print_case, rel
from oracle import adjoint as A
from oracle import les_cnn as L
from oracle import spectral as S

pytestmark = pytest.mark.gpu

N = 64
S_IN, S_OUT, EXTRA = 1.0, 0.5, 0.25
TOL_BF16 = 2e-2


def share_err(got, ref, ref0):
    return np.linalg.norm(got - ref) / np.linalg.norm(ref - ref0)


@pytest.fixture(scope="module")
def env():
    import symcode as sc

    go = S.Grid(3, N)
    gg = sc.Grid(3, N, dtype=torch.float64)
    net = sc.cnn.ConvClosureNet(N, batch_max=1, cin=4, width=64, cout=6, seed=3)
    h = net.get_params()
    h[-192:] = np.random.default_rng(4).standard_normal(192).astype(np.float32) * 0.05
    net.set_params(h)
    Ws, bs = net.unflatten(net.get_params())
    onet = L.Net([W.copy() for W in Ws], [b.copy() for b in bs], 4, S_IN, S_OUT, EXTRA, emulate=True)
    model = sc.create_learned_closure(net, gg, S_IN, S_OUT, EXTRA)
    u = oracle_solenoidal(go, 21, trunc=True)
    return sc, go, gg, net, onet, model, u


def grads_of(net):
    GW, GB = net.unflatten(net.get_grads())
    return GW, GB


def check_grads(net, og, tol=3e-2):
    GW, GB = grads_of(net)
    for l in range(4):
        assert rel(GW[l], og[0][l]) < tol, ("W", l, rel(GW[l], og[0][l]))
    for l in range(3):
        assert rel(GB[l], og[1][l]) < tol, ("b", l, rel(GB[l], og[1][l]))


def test_closure_tau(env):
    sc, go, gg, net, onet, model, u = env
    print_case("cnn closure tau", N)
    tau = gg.download_physical(model(gg.upload_spectral(u)))
    otau, _ = L.cnn_tau(u, go, onet)
    assert rel(tau, otau) < 1e-2


def test_les_rhs_closure4(env):
    sc, go, gg, net, onet, model, u = env
    print_case("les_rhs closure 4", N)
    nu = 0.01
    du = gg.download_spectral(sc.les_rhs(gg.vectorfield(), gg.upload_spectral(u), gg, model, nu))
    ref = L.les_rhs(u, go, nu, onet)
    ref0 = S.convectiondiffusion(u, go, nu)
    assert share_err(du, ref, ref0) < TOL_BF16


def test_les_rk4_closure4(env):
    sc, go, gg, net, onet, model, u = env
    print_case("les_rk4 closure 4", N)
    nu, dt = 0.01, 0.01
    ud = gg.upload_spectral(u)
    sc.les_rk4(ud, dt, gg, model, nu)
    ref = L.rk4(u, dt, go, nu, onet)
    ref0 = S.rk4(lambda v: S.convectiondiffusion(v, go, nu), u, dt, go)
    assert share_err(gg.download_spectral(ud), ref, ref0) < TOL_BF16


def test_les_rhs_vjp_closure4(env):
    sc, go, gg, net, onet, model, u = env
    print_case("les_rhs_vjp closure 4", N)
    nu = 0.01
    w = oracle_solenoidal(go, 22)
    net.zero_grad()
    out = sc.les_rhs_vjp(gg.vectorfield(), gg.upload_spectral(u), gg.upload_spectral(w), gg, model, nu)
    og = L.zero_grads(onet)
    ref = L.les_rhs_vjp(u, w, go, nu, onet, grads=og)
    ref0 = A.rhs_vjp(u, w, go, nu)
    assert share_err(gg.download_spectral(out), ref, ref0) < 3e-2
    check_grads(net, og)


def test_les_rk4_vjp_closure4(env):
    sc, go, gg, net, onet, model, u = env
    print_case("les_rk4_vjp closure 4", N)
    nu, dt = 0.01, 0.01
    ub = oracle_solenoidal(go, 23)
    net.zero_grad()
    out = sc.les_rk4_vjp(gg.upload_spectral(u), gg.upload_spectral(ub), dt, gg, model, nu)
    og = L.zero_grads(onet)
    ref = L.rk4_vjp(u, ub, dt, go, nu, onet, grads=og)
    ref0 = A.rk4_vjp(u, ub, dt, go, nu)
    assert share_err(gg.download_spectral(out), ref, ref0) < 3e-2
    check_grads(net, og)


@pytest.mark.parametrize("D,n", [(3, 16), (2, 32)])
def test_les_rk4_and_vjp_without_closure(D, n):
    """closure 0: the LES RK4 step and its transpose equal the DNS ones
    (band forcing on)"""
    import symcode as sc

    print_case("les_rk4 closure 0", D, n)
    go = S.Grid(D, n)
    gg = sc.Grid(D, n, dtype=torch.float64)
    sc.set_band_forcing(gg, 0.1, 2)
    u = oracle_solenoidal(go, 24, trunc=True)
    nu, dt = 0.01, 0.02
    f = lambda v: S.convectiondiffusion(v, go, nu, forcing=(0.1, 2))
    ud = gg.upload_spectral(u)
    sc.les_rk4(ud, dt, gg, None, nu)
    assert rel(gg.download_spectral(ud), S.rk4(f, u, dt, go)) < 1e-12
    ub = oracle_solenoidal(go, 25)
    out = sc.les_rk4_vjp(gg.upload_spectral(u), gg.upload_spectral(ub), dt, gg, None, nu)
    assert rel(gg.download_spectral(out), A.rk4_vjp(u, ub, dt, go, nu, (0.1, 2))) < 1e-12


def test_e_post_and_norm2():
    """e_post (reduce!, src/les.jl:94-97) and the unweighted half-spectrum
    norm against the oracle"""
    import symcode as sc
    from symcode import training as T
    from oracle import optim as O

    print_case("e_post", 3, 16)
    go = S.Grid(3, 16)
    gg = sc.Grid(3, 16, dtype=torch.float64)
    u = oracle_solenoidal(go, 31, trunc=True)
    r = oracle_solenoidal(go, 32)
    assert sc.solver.norm2(gg.upload_spectral(r), gg) == pytest.approx(np.linalg.norm(r.ravel()) ** 2, rel=1e-13)
    got = T.e_post(gg.upload_spectral(u), gg.upload_spectral(r), gg)
    ref = O.e_post(u, r, lambda c: S.twothirds(c, go))
    assert got == pytest.approx(ref, rel=1e-12)


def test_les_rk4_vjp_closure4_tracefree(env):
    """the learned closure with trace removal on the output: the RK4 step's
    transpose and the parameter gradient against the oracle"""
    sc, go, gg, net, onet, model, u = env
    print_case("les_rk4_vjp closure 4 tracefree", N)
    Ws, bs = net.unflatten(net.get_params())
    tnet = sc.cnn.ConvClosureNet(N, batch_max=1, cin=4, width=64, cout=6, seed=3, tracefree=True)
    tnet.set_params(net.get_params())
    tonet = L.Net([W.copy() for W in Ws], [b.copy() for b in bs], 4, S_IN, S_OUT, EXTRA, emulate=True,
                  tracefree=True)
    tmodel = sc.create_learned_closure(tnet, gg, S_IN, S_OUT, EXTRA)
    tau = gg.download_physical(tmodel(gg.upload_spectral(u)))
    otau, _ = L.cnn_tau(u, go, tonet)
    assert rel(tau, otau) < 1e-2
    assert np.abs(tau[0] + tau[1] + tau[2]).max() < 1e-5 * np.abs(tau[:3]).max()
    nu, dt = 0.01, 0.01
    ub = oracle_solenoidal(go, 26)
    tnet.zero_grad()
    out = sc.les_rk4_vjp(gg.upload_spectral(u), gg.upload_spectral(ub), dt, gg, tmodel, nu)
    og = L.zero_grads(tonet)
    ref = L.rk4_vjp(u, ub, dt, go, nu, tonet, grads=og)
    ref0 = A.rk4_vjp(u, ub, dt, go, nu)
    assert share_err(gg.download_spectral(out), ref, ref0) < 3e-2
    check_grads(tnet, og)


def test_rk4_vjp_kept_activations_equal_recompute(env):
    """With batch_max >= 4 the RK4 transpose keeps the network activations of
    its stage recomputes (slots 1-3) and starts three of the four stage
    transposes from them instead of evaluating the network again: the
    velocity cotangent and the weight gradients are the same bits as on the
    recomputing path (batch_max 1); the bias gradients (block sums added with
    atomics) agree to 1e-6."""
    sc, go, gg, net, onet, model, u = env
    print_case("rk4 vjp kept activations", N)
    net4 = sc.cnn.ConvClosureNet(N, batch_max=4, cin=4, width=64, cout=6, seed=3)
    net4.set_params(net.get_params())
    model4 = sc.create_learned_closure(net4, gg, S_IN, S_OUT, EXTRA)
    ud = gg.upload_spectral(u)
    wb = gg.upload_spectral(oracle_solenoidal(go, 22, trunc=True))
    res = []
    for nn, mm in ((net, model), (net4, model4)):
        nn.zero_grad()
        out = sc.les_rk4_vjp(ud, wb, 2e-3, gg, mm, 3e-3)
        res.append((gg.download_spectral(out), grads_of(nn)))
    model.attach()  # the module's closure stays attached for later tests
    (o1, (w1, b1)), (o4, (w4, b4)) = res
    assert np.array_equal(o1, o4)
    for l in range(4):
        assert np.array_equal(w1[l], w4[l]), l
    for l in range(3):
        assert rel(b4[l], b1[l]) < 1e-6, l
