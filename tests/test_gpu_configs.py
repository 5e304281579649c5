"""GPU tests on the workloads of the LES configurations: FP32 classical
closures (the 128^3 FP32 LES with dynamic Smagorinsky and Clark), the
512^3 -> 128^3 SFS pairs, filter_reynolds, and the FP32 a-posteriori
gradient through the learned closure (multi-step, against the oracle at 64^3,
and by central differences at 128^3).

Tolerances.  FP64 fields: 1e-12 relative L2.  FP32 fields against the FP64
oracle: 2e-5 for the right-hand sides and steps (single-precision FFTs and
products), 1e-4 for closure stresses (dynamic Smagorinsky's L is a difference
of two filtered products of similar size, which costs about a decade).  The
learned closure stores weights and activations in BF16: its share is compared
at 3e-2 as in the single-step tests."""
import numpy as np
import pytest
import torch

from helpers import oracle_solenoidal, print_case, rel
from oracle import closures as OC
from oracle import filtering as OF
from oracle import spectral as S

pytestmark = pytest.mark.gpu

TOL32 = 2e-5
TOL32_TAU = 1e-4


def grid_pair(D, n, prec):
    import symcode as sc

    return S.Grid(D, n), sc.Grid(D, n, dtype=torch.float64 if prec == 64 else torch.float32)


# ------------------------------------------------------------ filter_reynolds
def test_filter_reynolds_parity():
    """filter_reynolds truncates the gradient to the 2/3 box whatever ū holds:
    an untruncated ū against the oracle"""
    import symcode as sc
    from symcode import training as T

    print_case("filter_reynolds", 3, 16)
    go, gg = grid_pair(3, 16, 64)
    u = oracle_solenoidal(go, 41, trunc=False)
    ref = OF.filter_reynolds(u, go, 2e-3, 0.4)
    got = T.filter_reynolds(gg.upload_spectral(u), gg, 2e-3, 0.4)
    assert got == pytest.approx(ref, rel=1e-12)
    ut = np.stack([S.twothirds(c, go) for c in u])
    assert sc.filter_reynolds(gg.upload_spectral(ut), gg, 2e-3, 0.4) == pytest.approx(ref, rel=1e-12)


# ------------------------------------------------------- FP32 closures / LES
@pytest.mark.parametrize("kind", ["clark", "smag", "dynsmag"])
def test_fp32_closures_against_fp64_oracle(kind):
    """τ, les! and one RK4 LES step with FP32 fields at 32^3 against the FP64
    oracle, band forcing on"""
    import symcode as sc

    print_case("fp32 closure", kind, 32)
    n = 32
    go, gg = grid_pair(3, n, 32)
    u = oracle_solenoidal(go, 42, trunc=True)
    ud = gg.upload_spectral(u)
    delta = 2 * (2 * np.pi / n)
    nu = 0.01
    model, omodel = {
        "clark": (sc.create_clark(delta, gg), lambda u_, G: OC.clark(G, delta, 3)),
        "smag": (sc.create_smagorinsky(0.1, delta, gg), lambda u_, G: OC.smagorinsky(G, 0.1, delta, 3)),
        "dynsmag": (sc.create_dynamic_smagorinsky(delta, gg), lambda u_, G: OC.dynamic_smagorinsky(u_, delta, go)),
    }[kind]
    GG = sc.gradient_phys(ud, gg)
    oGG = OC.gradient_phys(u, go)
    assert rel(gg.download_physical(GG), oGG) < TOL32
    tau = gg.download_physical(model(ud, GG))
    e_tau = rel(tau, omodel(u, oGG))
    if kind == "dynsmag":
        _, oc = OC.dynamic_smagorinsky(u, delta, go, return_coeff=True)
        e_c = abs(model.last_coefficient - oc) / oc
        print(f"dynsmag c fp32 {model.last_coefficient:.8e} oracle {oc:.8e} rel {e_c:.2e}")
        assert e_c < TOL32_TAU
    print(f"{kind}: tau rel {e_tau:.2e}")
    assert e_tau < TOL32_TAU
    sc.set_band_forcing(gg, 0.1, 2)
    try:
        du = sc.les_rhs(gg.vectorfield(), ud, gg, model, nu)
        ref = OC.les_rhs(u, go, nu, omodel) + S.band_forcing(u, go, 0.1, 2)
        e_rhs = rel(gg.download_spectral(du), ref)
        v = ud.clone()
        sc.les_rk4(v, 0.01, gg, model, nu)
        f = lambda w: OC.les_rhs(w, go, nu, omodel) + S.band_forcing(w, go, 0.1, 2)
        e_step = rel(gg.download_spectral(v), S.rk4(f, u, 0.01, go))
    finally:
        sc.set_band_forcing(gg, 0.0)
    print(f"{kind}: les_rhs rel {e_rhs:.2e}, rk4 step rel {e_step:.2e}")
    assert e_rhs < TOL32 and e_step < TOL32


def _les_setup_128(prec):
    """a developed 128^3 field: the random-phase initial field (on which the
    dynamic coefficient clips to zero: it has no cascade yet) advanced by 60
    forced RK4 steps without a closure"""
    import symcode as sc

    g = sc.Grid(3, 128, dtype=torch.float64 if prec == 64 else torch.float32)
    u = sc.randomfield(sc.peak_profile, g, seed=7, totalenergy=0.2, kpeak=4)
    sc.set_band_forcing(g, 0.1, 2)
    for _ in range(60):
        sc.rk4_cfl(u, g, 2e-3, 0.3)
    sc.set_band_forcing(g, 0.0)
    return sc, g, u


@pytest.mark.parametrize("kind", ["clark", "dynsmag"])
def test_les_128_fp32_budget_and_fp64(kind):
    """Full size (128^3 FP32, the configuration's LES): (a) the closure's
    energy exchange identity Re<u, les!(u) - les!_none(u)> = <τ : ∇u> (the
    spectral inner product of a truncated u equals the physical mean), with
    both sides computed on the device in FP32; (b) 20 RK4 LES steps in FP32
    against the same steps in FP64 on the device (the FP64 path is pinned to
    the oracle at 1e-12): shell spectra within 1e-4, energy within 1e-5.
    The identity's FP32 tolerance: 1e-4 for dynamic Smagorinsky (measured
    1e-7), 1e-3 for Clark, whose τ = Δ²/12 G Gᵀ is not trace-free: the
    gradient of its trace part is orthogonal to the solenoidal u, and its
    contraction with tr S = ∇·u vanishes, only to FP32 rounding on either
    side, which the difference of the two inner products magnifies (measured
    4e-4)."""
    import symcode as sc
    from symcode import analysis as AN
    from symcode import training as T

    print_case("les 128 fp32", kind)
    nu = 2e-3
    delta = 2 * (2 * np.pi / 128)
    res = {}
    for prec in (64, 32):
        sc_, g, u = _les_setup_128(prec)
        if prec == 32:  # same initial field, rounded to FP32
            g64 = res[64]["g"]
            u = T.regrid(res[64]["u0"], g64, g)
        model = sc.create_clark(delta, g) if kind == "clark" else sc.create_dynamic_smagorinsky(delta, g)
        sc.set_band_forcing(g, 0.1, 2)
        if prec == 32:
            d1 = sc.les_rhs(g.vectorfield(), u, g, model, nu)
            d0 = sc.les_rhs(g.vectorfield(), u, g, None, nu)
            lhs = sc.spectraldot(u, d1, g).real - sc.spectraldot(u, d0, g).real
            GG = sc.gradient_phys(u, g)
            tau = model(u, GG)
            if kind == "dynsmag":
                print(f"dynsmag c = {model.last_coefficient:.6e}")
                assert model.last_coefficient > 0
            rhs = -AN.sample_mean(AN.sfs_dissipation(tau, GG, g), g)
            print(f"{kind} fp32 128: exchange {lhs:.8e} vs <tau:grad u> {rhs:.8e} rel {abs(lhs - rhs) / abs(rhs):.2e}")
            assert abs(lhs - rhs) < (1e-3 if kind == "clark" else 1e-4) * abs(rhs)
        dt = 0.3 * sc.propose_timestep(u, g, nu) if prec == 64 else res[64]["dt"]
        u0 = u.clone()
        for _ in range(20):
            sc.les_rk4(u, dt, g, model, nu)
        _, spec = sc.spectrum(u, g)
        res[prec] = {"g": g, "u0": u0, "dt": dt, "spec": np.asarray(spec), "e": sc.energy(u, g)}
        sc.set_band_forcing(g, 0.0)
    s64, s32 = res[64]["spec"], res[32]["spec"]
    es = np.linalg.norm(s32 - s64) / np.linalg.norm(s64)
    ee = abs(res[32]["e"] - res[64]["e"]) / res[64]["e"]
    print(f"{kind} 128: FP32 vs FP64 after 20 steps: spectrum {es:.2e}, energy {ee:.2e}")
    assert es < 1e-4 and ee < 1e-5


# ---------------------------------------------------- sfs 512^3 -> 128^3
def test_sfs_512_to_128_closed_form():
    """sfs! at the configuration's sizes on a band-limited field with an
    analytic answer: u = Re(a e^{ik.x}), k = (3, 1, 2) (|k|^2 = 14) on the
    512^3 DNS grid.  ū = g(k) u and τ = tracefree((g(2k) - g(k)^2)/2
    Re(a_i a_j e^{2ik.x})) on the 128^3 grid, g the Gaussian's transfer."""
    import symcode as sc

    print_case("sfs 512 -> 128")
    gd = sc.Grid(3, 512, dtype=torch.float64)
    gl = sc.Grid(3, 128, dtype=torch.float64)
    k = (3, 1, 2)
    a = np.array([0.7, -0.3 + 0.4j, 0.2j])
    u = gd.vectorfield()
    # rfft coefficient of Re(a e^{ik.x}) at +k is a / 2 (kx > 0: no conjugate partner stored)
    for c in range(3):
        u[c, k[2], k[1], k[0]] = complex(a[c]) / 2
    delta = 2 * (2 * np.pi / 128)
    ubar, tau = sc.sfs(u, gd, gl, delta)
    go = S.Grid(3, 128)
    n = 128
    x = np.arange(n) * (2 * np.pi / n)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    ph = k[0] * X + k[1] * Y + k[2] * Z
    k2 = float(sum(q * q for q in k))
    g1 = np.exp(-delta**2 * k2 / 24)
    g2 = np.exp(-delta**2 * 4 * k2 / 24)
    ub_ref = np.stack([S.to_spec(np.real(g1 * a[c] * np.exp(1j * ph)), go) for c in range(3)])
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0)]
    T = np.stack([(g2 - g1 * g1) / 2 * np.real(a[i] * a[j] * np.exp(2j * ph)) for i, j in pairs])
    T = OF.make_tracefree(T, 3)
    tau_ref = np.stack([S.to_spec(t, go) for t in T])
    e_u = rel(gl.download_spectral(ubar), ub_ref)
    e_t = rel(gl.download_spectral(tau), tau_ref)
    print(f"sfs 512->128: ubar rel {e_u:.2e}, tau rel {e_t:.2e}")
    assert e_u < 1e-12 and e_t < 1e-12


# ----------------------------------------- a-posteriori gradient, learned closure
def _closure_net(n, seed, scale_last=0.05):
    import symcode as sc

    net = sc.cnn.ConvClosureNet(n, batch_max=1, cin=4, width=64, cout=6, seed=seed)
    h = net.get_params()
    h[-192:] = np.random.default_rng(seed + 1).standard_normal(192).astype(np.float32) * scale_last
    net.set_params(h)
    return net


def test_rollout_gradient_fp32_multistep_against_oracle():
    """Two fixed-dt RK4 steps of the FP32 64^3 LES with the learned closure:
    loss, dL/du0 and the parameter gradient against oracle/les_cnn.py
    (BF16 storage points emulated).  dL/du0 is compared on the closure's
    share (oracle with closure minus oracle with a zero closure)."""
    import symcode as sc
    from oracle import les_cnn as L
    from symcode import training as T

    print_case("rollout gradient fp32 64 x2")
    n, nu, dt = 64, 0.01, 0.01
    s_in, s_out, extra = 1.0, 0.5, 0.25
    go = S.Grid(3, n)
    g = sc.Grid(3, n, dtype=torch.float32)
    net = _closure_net(n, 5)
    Ws, bs = net.unflatten(net.get_params())
    onet = L.Net([W.copy() for W in Ws], [b.copy() for b in bs], 4, s_in, s_out, extra, emulate=True)
    znet = L.Net([W.copy() for W in Ws[:3]] + [0 * Ws[3]], [b.copy() for b in bs], 4, s_in, s_out, extra,
                 emulate=True)
    model = sc.create_learned_closure(net, g, s_in, s_out, extra)
    u0 = oracle_solenoidal(go, 43, trunc=True)
    refs = [0.9 * oracle_solenoidal(go, 44 + t, trunc=True) for t in range(2)]
    net.zero_grad()
    Lg, gu = T.rollout_gradient(g.upload_spectral(u0), [g.upload_spectral(r) for r in refs], dt, g, nu, model)
    Lo, go_u, ogr = L.rollout_gradient(u0, refs, dt, go, nu, onet)
    _, g0_u, _ = L.rollout_gradient(u0, refs, dt, go, nu, znet)
    e_L = abs(Lg - Lo) / abs(Lo)
    e_u = np.linalg.norm(g.download_spectral(gu) - go_u) / np.linalg.norm(go_u - g0_u)
    GW, GB = net.unflatten(net.get_grads())
    e_w = [rel(GW[l], ogr[0][l]) for l in range(4)]
    e_b = [rel(GB[l], ogr[1][l]) for l in range(3)]
    print(f"rollout gradient: L rel {e_L:.2e}, dL/du0 share {e_u:.2e}, W {np.round(e_w, 4)}, b {np.round(e_b, 4)}")
    assert e_L < 1e-4
    assert e_u < 3e-2
    assert max(e_w) < 3e-2 and max(e_b) < 3e-2


def bf16_round(a):
    """float32 -> nearest bfloat16 (ties to even), returned as float32"""
    b = np.ascontiguousarray(a, np.float32).view(np.uint32)
    return ((b + (((b >> 16) & 1) + 0x7FFF)) & 0xFFFF0000).astype(np.uint32).view(np.float32)


def bf16_ulp(v):
    """spacing of bfloat16 at the (bfloat16) values v; 0 where v == 0"""
    out = np.zeros(v.shape, np.float64)
    nz = v != 0
    out[nz] = 2.0 ** (np.floor(np.log2(np.abs(v[nz].astype(np.float64)))) - 7)
    return out


def test_rollout_gradient_fd_128_fp32():
    """The configuration's gradient check on the GPU path at 128^3 FP32:
    central differences of the 10-step rollout loss along 8 random weight
    directions, against the adjoint's directional derivative.  The network
    stores its weights in BF16, so a perturbation smaller than one BF16 step
    of a weight changes the network only through rounding flips, whose noise
    is as large as the directional derivative of a dense random direction.
    The weights are therefore set to BF16 values and each direction moves
    every weight by exactly one BF16 step, with a random sign (eps d_i =
    ±ulp(w_i), eps = 1e-3): the perturbed networks are exactly the ones the
    MFMA kernels run, and what remains is the activations' BF16 rounding,
    which averages out over the 128^3 voxels.  Agreement within 2 % over the
    8 directions and 5 % on each."""
    import symcode as sc
    from symcode import training as T

    print_case("rollout gradient fd 128 fp32")
    n, nu = 128, 2e-3
    g = sc.Grid(3, n, dtype=torch.float32)
    u0 = sc.randomfield(sc.peak_profile, g, seed=8, totalenergy=0.2, kpeak=4)
    delta = 2 * (2 * np.pi / n)
    dt = 0.3 * sc.propose_timestep(u0, g, nu)
    refs, r = [], u0.clone()
    for _ in range(10):
        sc.les_rk4(r, dt, g, None, nu)
        refs.append(r.clone())
    s_in = T.velocity_rms(u0, g)
    s_out = delta**2 * sc.get_dissipation(u0, 1.0, g)
    net = _closure_net(n, 9, scale_last=0.5)
    model = sc.create_learned_closure(net, g, s_in, s_out, 0.1)
    nw = sum(27 * co * ci for co, ci in net.layer_shapes())
    theta = net.get_params()
    theta[:nw] = bf16_round(theta[:nw])
    net.set_params(theta)
    net.zero_grad()
    L0, _ = T.rollout_gradient(u0, refs, dt, g, nu, model)
    grad = net.get_grads().astype(np.float64)
    # padding entries (input channels 4..7 of layer 1, outputs 6..15 of layer 4) stay zero
    mask = np.ones(theta.size, np.float32)
    MW, _ = net.unflatten(mask)
    MW[0][:, :, net.cin:] = 0
    MW[3][:, net.cout:, :] = 0
    mask[nw:] = 0  # biases (FP32 in the kernels) are not moved
    ulp = bf16_ulp(theta) * mask
    rng = np.random.default_rng(10)
    fd, an = [], []
    for _ in range(8):
        step = rng.choice([-1.0, 1.0], theta.size) * ulp  # eps * d
        net.set_params((theta + step).astype(np.float32))
        Lp = T.rollout_loss(u0, refs, dt, g, nu, model)
        net.set_params((theta - step).astype(np.float32))
        Lm = T.rollout_loss(u0, refs, dt, g, nu, model)
        fd.append((Lp - Lm) / 2)
        an.append(float(grad @ step))
    net.set_params(theta)
    fd, an = np.array(fd), np.array(an)
    per = np.abs(fd - an) / np.abs(an)
    tot = np.linalg.norm(fd - an) / np.linalg.norm(an)
    print(f"fd 128: L0 {L0:.6e}\n  fd {fd}\n  an {an}\n  per-direction rel {per}\n  overall {tot:.3e}")
    assert tot < 2e-2 and per.max() < 5e-2
