"""What the trained network gains a-posteriori, over the LES without a closure
and over dynamic Smagorinsky.

A 256³ forced DNS (FP64, ν = 2e-3, band forcing 0.1 on |k| <= 2) is spun up
to t = 1 and filtered to 64³ (Gaussian Δ = 2 h_LES, sfs!) at twelve save
times 0.05 apart; the convolutional closure is trained on those pairs with
the reference's semantics (trace-free output, relative loss, ClipNorm(1) +
AdamW, warm-up + cosine; src/training.jl:3-9,45-70).  Then 64³ FP32 LES
(RK4, fixed dt, band forcing) run 100 steps from the filtered DNS of the
next time, and each state is compared with the filtered DNS advanced to the
same time (the DNS takes its own CFL sub-steps): the shell-spectrum error and
e_post (reduce!, src/les.jl:94-97) after 100 steps and the trajectory MSE
over the 100.  Measured on MI355X: no closure 0.0229 / 0.126 / 3.67e-3,
dynamic Smagorinsky 0.0147 / 0.092 / 2.02e-3, the trained network 0.0034 /
0.069 / 1.15e-3 (Clark 0.0019 / 0.054 / 6.7e-4)."""
import numpy as np
import pytest
import torch

from helpers import print_case

pytestmark = pytest.mark.gpu


def spectrum_error(sc, a, b, g):
    _, sa = sc.spectrum(a, g)
    _, sb = sc.spectrum(b, g)
    m = g.n // 3
    sa, sb = np.asarray(sa[:m]), np.asarray(sb[:m])
    return float(np.linalg.norm(sa - sb) / np.linalg.norm(sb))


def test_aposteriori_gain_over_no_closure_and_dynsmag():
    import symcode as sc
    from symcode import training as T

    print_case("aposteriori gain 256 -> 64")
    visc, cfl = 2e-3, 0.3
    g_dns = sc.Grid(3, 256, dtype=torch.float64)
    g_les = sc.Grid(3, 64, dtype=torch.float64)
    g32 = sc.Grid(3, 64, dtype=torch.float32)
    u = sc.randomfield(sc.peak_profile, g_dns, seed=11, totalenergy=0.2, kpeak=4)
    sc.set_band_forcing(g_dns, 0.1, 2)
    T.create_data(u, g_dns, g_les, visc, [2], 1, 1.0, cfl=cfl)  # spin-up to t = 1
    pairs = T.create_data(u, g_dns, g_les, visc, [2], 12, 0.05, cfl=cfl)[2]
    delta = pairs[0].delta
    nrm = T.Normalization.of(pairs, g_les)
    net = sc.cnn.ConvClosureNet(64, batch_max=12, cin=4, width=64, cout=6, seed=0, tracefree=True)
    x, y = T.pack_pairs(net, g_les, pairs, nrm)
    losses = T.reference_train(net, x, y, 400, lr=1e-3)
    print(f"relative loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
    assert losses[-1] < 0.1 * losses[0]
    sc.set_band_forcing(g32, 0.1, 2)
    u0 = T.regrid(sc.sfs(u, g_dns, g_les, delta)[0], g_les, g32)
    dt = 0.5 * cfl * sc.propose_timestep(u, g_dns, visc) * (256 / 64)
    refs = []
    for _ in range(100):
        tt = 0.0
        while tt < dt * (1 - 1e-12):
            tt += sc.rk4_cfl(u, g_dns, visc, cfl, dt - tt)
        refs.append(T.regrid(sc.sfs(u, g_dns, g_les, delta)[0], g_les, g32))
    models = {
        "none": None,
        "dynsmag": sc.create_dynamic_smagorinsky(delta, g32),
        "cnn": nrm.closure(net, g32, delta, pairs[-1].redelta),
    }
    res = {}
    e = g32.vectorfield()
    for name, m in models.items():
        v = u0.clone()
        traj = 0.0
        for r in refs:
            sc.les_rk4(v, dt, g32, m, visc)
            sc.lincomb(e, v, 1.0, r, -1.0, g32)
            traj += sc.spectraldot(e, e, g32).real
        res[name] = (spectrum_error(sc, v, refs[-1], g32), T.e_post(v, refs[-1], g32), traj / len(refs))
        print(f"{name}: spectrum error {res[name][0]:.5f}, e_post {res[name][1]:.5f}, "
              f"trajectory MSE {res[name][2]:.4e}", flush=True)
    sc.set_band_forcing(g32, 0.0)
    sc.set_band_forcing(g_dns, 0.0)
    cnn, none, dyn = res["cnn"], res["none"], res["dynsmag"]
    # against the LES without a closure: every measure, by a wide margin
    assert cnn[0] < 0.5 * none[0] and cnn[1] < 0.8 * none[1] and cnn[2] < 0.6 * none[2]
    # against dynamic Smagorinsky: every measure
    assert cnn[0] < 0.5 * dyn[0] and cnn[1] < 0.9 * dyn[1] and cnn[2] < 0.8 * dyn[2]
