"""The learned closure trained on the GPU DNS's own filtered pairs, and what it
gains over the closures without it.

A 128³ forced DNS (FP64) is filtered to 64³ (Gaussian Δ = 2 h_LES, sfs!)
at six save times; the network is trained on the first five and judged on
the sixth (held out, later in time) by the a-priori metrics of
compute_sfs_stats (src/analysis.jl:116-121): relative error and correlation
of τ.  No model has relative error 1 and no correlation; the test asks the
trained network to beat that and dynamic Smagorinsky's correlation."""
import math

import numpy as np
import pytest
import torch

from helpers import print_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def data():
    import symcode as sc
    from symcode import training as T

    g_dns = sc.Grid(3, 128, dtype=torch.float64)
    g_les = sc.Grid(3, 64, dtype=torch.float64)
    visc = 2e-3
    u = sc.randomfield(lambda k: sc.peak_profile(k, 4.0), g_dns, seed=11, totalenergy=1.0)
    sc.set_band_forcing(g_dns, 0.1, 2)
    T.create_data(u, g_dns, g_les, visc, [2], 1, 0.5, cfl=0.3)  # spin-up to t = 0.5
    pairs = T.create_data(u, g_dns, g_les, visc, [2], 6, 0.1, cfl=0.3)[2]
    return sc, T, g_les, visc, pairs


def test_apriori_gain(data):
    sc, T, g, visc, pairs = data
    print_case("apriori gain 128->64")
    train, test = pairs[:5], pairs[5]
    delta = test.delta
    nrm = T.Normalization.of(train, g)
    net = sc.cnn.ConvClosureNet(64, batch_max=5, cin=4, width=64, cout=6, seed=0)
    x, y = T.pack_pairs(net, g, train, nrm)
    losses = T.train(net, x, y, 400, lr=1e-3)
    print("loss", losses[0], losses[-1], flush=True)
    assert losses[-1] < 0.5 * losses[0]
    tau_ref = sc.to_phys(g.spacetensorfield(), test.tau, g)
    model = nrm.closure(net, g, delta, test.redelta)
    tau_cnn = model(test.ubar)
    m_cnn = T.apriori_metrics(tau_cnn, tau_ref)
    tau_dyn = sc.create_dynamic_smagorinsky(delta, g)(test.ubar)
    m_dyn = T.apriori_metrics(tau_dyn, tau_ref)
    GG = sc.gradient_phys(test.ubar, g)
    tau_clark = sc.make_tracefree(sc.create_clark(delta, g)(test.ubar, GG), g, spectral=False)
    m_clark = T.apriori_metrics(tau_clark, tau_ref)
    print("relerr/crosscor  cnn", m_cnn, " dynsmag", m_dyn, " clark", m_clark, flush=True)
    assert m_cnn[0] < 0.9
    assert m_cnn[1] > m_dyn[1]


def test_apriori_gain_reference_semantics(data):
    """the same gain with the reference's training semantics: trace-free
    output, relative loss, ClipNorm(1) + AdamW, 5 % warm-up then cosine decay
    (src/training.jl:3-9,45-70; src/nets.jl:275-293)"""
    sc, T, g, visc, pairs = data
    print_case("apriori gain reference semantics 128->64")
    train, test = pairs[:5], pairs[5]
    nrm = T.Normalization.of(train, g)
    net = sc.cnn.ConvClosureNet(64, batch_max=5, cin=4, width=64, cout=6, seed=0, tracefree=True)
    x, y = T.pack_pairs(net, g, train, nrm)
    losses = T.reference_train(net, x, y, 400, lr=1e-3)
    print("relative loss", losses[0], losses[-1], flush=True)
    assert losses[-1] < 0.5 * losses[0]
    tau_ref = sc.to_phys(g.spacetensorfield(), test.tau, g)
    tau_cnn = nrm.closure(net, g, test.delta, test.redelta)(test.ubar)
    tc = tau_cnn.double()
    assert float((tc[0] + tc[1] + tc[2]).abs().max()) < 1e-5 * float(tc[:3].abs().max())
    m_cnn = T.apriori_metrics(tau_cnn, tau_ref)
    m_dyn = T.apriori_metrics(sc.create_dynamic_smagorinsky(test.delta, g)(test.ubar), tau_ref)
    print("relerr/crosscor  cnn", m_cnn, " dynsmag", m_dyn, flush=True)
    assert m_cnn[0] < 0.9
    assert m_cnn[1] > m_dyn[1]


def test_aposteriori_gradient_step_reduces_rollout_loss(data):
    """a few a-posteriori Adam steps on a 64³ FP64 LES rollout (8 RK4 steps)
    against the filtered DNS lower its loss"""
    sc, T, g, visc, pairs = data
    print_case("aposteriori steps 64")
    nrm = T.Normalization.of(pairs, g)
    net = sc.cnn.ConvClosureNet(64, batch_max=1, cin=4, width=64, cout=6, seed=1)
    x, y = T.pack_pairs(net, g, pairs[:1], nrm)
    T.train(net, x, y, 100, lr=1e-3)
    model = nrm.closure(net, g, pairs[0].delta, pairs[0].redelta)
    # reference trajectory: the six save times are 0.1 apart; 4 RK4 steps per interval
    dt = 0.1 / 4
    u0 = pairs[0].ubar
    # the filtered DNS is saved every 4th LES step; the states between are
    # linear interpolations of the saved fields
    full = []
    prev = u0
    for k in range(2):
        nxt = pairs[k + 1].ubar
        for j in range(1, 5):
            r = g.vectorfield()
            sc.lincomb(r, prev, 1 - j / 4, nxt, j / 4, g)
            full.append(r)
        prev = nxt
    L0 = None
    for it in range(4):
        L = T.aposteriori_step(net, u0, full, dt, g, visc, model, lr=3e-4)
        print("aposteriori loss", it, L, flush=True)
        assert math.isfinite(L)
        L0 = L if L0 is None else L0
    assert L < L0
    # a-posteriori error of the final state (e_post, src/les.jl:97)
    u = u0.clone()
    for _ in full:
        sc.les_rk4(u, dt, g, model, visc)
    e = T.e_post(u, full[-1], g)
    print("e_post after", len(full), "steps", e, flush=True)
    assert 0 < e < 1
