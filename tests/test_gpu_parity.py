"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

Bar: FP64 within 1e-12 relative (L2 over the field) per call / per step;
FP32 within the stated single-precision tolerances."""
import numpy as np
import pytest
import torch

from helpers import oracle_solenoidal, print_case, rand_phys, rel
from oracle import closures as OC
from oracle import filtering as OF
from oracle import spectral as S

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 2e-5

SIZES3 = [8, 12, 16, 30, 32, 48, 64, 96]
SIZES2 = [16, 64, 256, 810]


def grids(D, n, prec=64):
    import symcode as sc

    dt = torch.float64 if prec == 64 else torch.float32
    return S.Grid(D, n, 2 * np.pi, np.float64), sc.Grid(D, n, 2 * np.pi, dt)


def cases():
    return [(3, n) for n in SIZES3] + [(2, n) for n in SIZES2]


# ---------------------------------------------------------------- transforms
@pytest.mark.parametrize("D,n", cases())
def test_fft_parity(D, n):
    import symcode as sc

    print_case("fft", D, n)
    go, gg = grids(D, n)
    x = rand_phys((2,) + go.phys_shape, 1)
    ref = np.stack([S.to_spec(x[i], go) for i in range(2)])
    spec = sc.to_spec(gg.spec(2), gg.upload_physical(x), gg)
    assert rel(gg.download_spectral(spec), ref) < TOL64
    back = sc.to_phys(gg.phys(2), spec, gg)
    assert rel(gg.download_physical(back), x) < TOL64
    # C2R of a non-Hermitian spectrum follows the multi-dimensional C2R
    # convention (imaginary parts of kx = 0 and kx = n/2 discarded)
    r = np.random.default_rng(2)
    z = r.standard_normal(go.spec_shape) + 1j * r.standard_normal(go.spec_shape)
    p = sc.to_phys(gg.phys(1), gg.upload_spectral(z), gg)
    assert rel(gg.download_physical(p)[0], S.to_phys(z, go)) < TOL64
    # plan-level scalings: ldiv! normalised, mul! unnormalised
    q = sc.irfft(gg.phys(1), gg.upload_spectral(ref[0]), gg)
    assert rel(gg.download_physical(q)[0], S.irfft_normalised(ref[0], go)) < TOL64
    s2 = sc.rfft(gg.spec(1), gg.upload_physical(x[0]), gg)
    assert rel(gg.download_spectral(s2)[0], S.rfft_unnormalised(x[0], go)) < TOL64


# ----------------------------------------------------------- spectral kernels
@pytest.mark.parametrize("D,n", [(3, 12), (3, 30), (2, 64)])
def test_pointwise_kernels(D, n):
    import symcode as sc

    print_case("kernels", D, n)
    go, gg = grids(D, n)
    u = oracle_solenoidal(go, 3) * 1.3 + 0.1 * np.stack([S.to_spec(rand_phys(go.phys_shape, 9 + i), go) for i in range(D)])
    ud = gg.upload_spectral(u)
    t = gg.upload_spectral(u)
    sc.twothirds(t, gg)
    assert np.array_equal(gg.download_spectral(t), np.stack([S.twothirds(c, go) for c in u]))
    t = gg.upload_spectral(u)
    sc.project(t, gg)
    assert rel(gg.download_spectral(t), S.project(u, go)) < TOL64
    sig = np.stack([S.to_spec(rand_phys(go.phys_shape, 20 + i), go) for i in range(go.tensordim)])
    t = gg.upload_spectral(sig)
    sc.viscosity(t, ud, 0.03, gg)
    assert rel(gg.download_spectral(t), S.viscosity(sig, u, 0.03, go)) < TOL64
    du = sc.tensordivergence(gg.vectorfield(), gg.upload_spectral(sig), gg)
    assert rel(gg.download_spectral(du), S.tensordivergence(sig, go)) < TOL64
    G = sc.vectorgradient(gg.tensorfield_nonsym(), ud, gg)
    assert rel(gg.download_spectral(G), S.vectorgradient(u, go)[0]) < TOL64
    St = sc.strainrate(gg.tensorfield(), ud, gg)
    assert rel(gg.download_spectral(St), S.strainrate(u, go)) < TOL64
    dv = sc.vectordivergence(gg.scalarfield(), ud, gg)
    assert rel(gg.download_spectral(dv)[0], S.vectordivergence(u, go)) < TOL64
    t = gg.upload_spectral(u)
    sc.gaussianfilter(t, 0.37, gg)
    assert rel(gg.download_spectral(t), np.stack([OF.gaussianfilter(c, 0.37, go) for c in u])) < TOL64
    # symmetrize on raw normals
    z = S.random_spectral(go, 4, 1)[0]
    t = gg.upload_spectral(z)
    sc.symmetrize(t, gg)
    assert np.array_equal(gg.download_spectral(t)[0], S.symmetrize(z, go))
    # axpy and trace removal
    t = gg.upload_spectral(sig)
    sc.filtering.axpy(t, gg.upload_spectral(sig[::-1].copy()), -0.5, gg)
    assert rel(gg.download_spectral(t), sig - 0.5 * sig[::-1]) < TOL64
    t = gg.upload_spectral(sig)
    sc.make_tracefree(t, gg)
    assert rel(gg.download_spectral(t), OF.make_tracefree(sig, D)) < TOL64


# --------------------------------------------------------------------- RHS
RHS_CASES = [(3, 12), (3, 16), (3, 30), (3, 32), (3, 48), (2, 32), (2, 256), (2, 810)]


@pytest.mark.parametrize("D,n", RHS_CASES)
def test_rhs_parity(D, n):
    import symcode as sc

    print_case("rhs", D, n)
    go, gg = grids(D, n)
    u = oracle_solenoidal(go, 5)  # untruncated: exercises the full viscous path
    nu = 0.02
    ud = gg.upload_spectral(u)
    sig = sc.nonlinearity(gg.tensorfield(), ud, gg)
    assert rel(gg.download_spectral(sig), S.nonlinearity(u, go)) < TOL64
    st = sc.stress(gg.tensorfield(), ud, nu, gg)
    assert rel(gg.download_spectral(st), S.stress(u, nu, go)) < TOL64
    du = sc.convectiondiffusion(gg.vectorfield(), ud, gg, nu)
    assert rel(gg.download_spectral(du), S.convectiondiffusion(u, go, nu)) < TOL64


@pytest.mark.parametrize("D,n", [(3, 16), (3, 30), (2, 64)])
def test_steppers_parity(D, n):
    import symcode as sc

    print_case("steppers", D, n)
    go, gg = grids(D, n)
    u = oracle_solenoidal(go, 6, trunc=True)
    nu, dt = 0.01, 0.02
    f = lambda v: S.convectiondiffusion(v, go, nu)
    for name, ostep, gstep in (("wray3", S.wray3, sc.wray3), ("rk4", S.rk4, sc.rk4)):
        ref = u
        ud = gg.upload_spectral(u)
        for k in range(3):
            ref = ostep(f, ref, dt, go)
            gstep(ud, dt, gg, nu)
            assert rel(gg.download_spectral(ud), ref) < TOL64 * (k + 1), (name, k)


@pytest.mark.parametrize("D,n", [(3, 16), (2, 64)])
def test_band_forcing_parity(D, n):
    import symcode as sc

    print_case("band", D, n)
    go, gg = grids(D, n)
    u = oracle_solenoidal(go, 8, trunc=True)
    sc.set_band_forcing(gg, 0.1, 2)
    du = sc.convectiondiffusion(gg.vectorfield(), gg.upload_spectral(u), gg, 0.01)
    ref = S.convectiondiffusion(u, go, 0.01, forcing=(0.1, 2))
    assert rel(gg.download_spectral(du), ref) < TOL64
    # injection by the forcing alone is exactly the requested rate
    f = gg.download_spectral(du) - S.convectiondiffusion(u, go, 0.01)
    inj = sum(S.spectraldot(u[i], f[i], go) for i in range(D)).real
    assert inj == pytest.approx(0.1, rel=1e-11)
    sc.set_band_forcing(gg, 0.0)


@pytest.mark.parametrize("D,n", [(3, 16), (3, 30), (2, 64)])
def test_reductions_and_cfl(D, n):
    import symcode as sc

    print_case("reductions", D, n)
    go, gg = grids(D, n)
    u = oracle_solenoidal(go, 9)
    ud = gg.upload_spectral(u)
    assert sc.energy(ud, gg) == pytest.approx(S.energy(u, go), rel=1e-13)
    assert sc.get_dissipation(ud, 0.01, gg) == pytest.approx(S.get_dissipation(u, 0.01, go), rel=1e-13)
    v = gg.upload_spectral(u[::-1].copy())
    dot = sc.spectraldot(ud, v, gg)
    ref = sum(S.spectraldot(u[i], u[::-1][i], go) for i in range(D))
    assert abs(dot - ref) < 1e-13 * abs(ref)
    k, s = sc.spectrum(ud, gg)
    ko, so = S.spectrum(u, go)
    assert np.allclose(k, ko) and rel(s, so) < 1e-13
    assert sc.propose_timestep(ud, gg, 0.01) == pytest.approx(S.propose_timestep(u, go, 0.01), rel=1e-12)
    shells = sc.energy_shells(gg, [1, 2], ud)
    oshells = S.energy_shells(u, go)
    for (i, e), (j, eo) in zip(shells, oshells):
        assert i == j and e == pytest.approx(eo, rel=1e-13)
    w = gg.upload_spectral(1.5 * u)
    sc.maintain_shell_energy(w, shells, gg)
    assert rel(gg.download_spectral(w), S.maintain_shell_energy(1.5 * u, go, oshells)) < TOL64
    stats = sc.turbulence_statistics(gg.upload_spectral(u), 0.01, gg)
    _, ostats = S.turbulence_statistics(u, 0.01, go)
    for key in ostats:
        assert stats[key] == pytest.approx(ostats[key], rel=1e-12), key


@pytest.mark.parametrize("D,n", [(3, 16), (3, 32), (2, 64)])
def test_initial_conditions(D, n):
    import symcode as sc

    print_case("ic", D, n)
    go, gg = grids(D, n)
    prof = lambda k: S.peak_profile(k, 4)
    u = sc.randomfield(lambda k: prof(k), gg, seed=3, totalenergy=1.0)
    ref = S.randomfield(prof, go, seed=3, totalenergy=1.0)
    assert rel(gg.download_spectral(u), ref) < TOL64
    tg = sc.taylorgreen(gg, V0=1.3)
    assert rel(gg.download_spectral(tg), S.taylorgreen(go, V0=1.3)) < TOL64


# ---------------------------------------------------------------- filtering
@pytest.mark.parametrize("nd,nl,D", [(48, 16, 3), (32, 16, 3), (256, 64, 2)])
def test_cutoff_and_sfs(nd, nl, D):
    import symcode as sc

    print_case("sfs", nd, nl, D)
    god, ggd = grids(D, nd)
    gol, ggl = grids(D, nl)
    u = oracle_solenoidal(god, 10, trunc=True)
    ud = ggd.upload_spectral(u)
    ub = sc.cutoff(ggl.vectorfield(), ud, ggl, ggd)
    assert np.array_equal(ggl.download_spectral(ub), np.stack([OF.cutoff(c, gol, god) for c in u]))
    delta = 2 * (2 * np.pi / nl)
    ubar, tau = sc.sfs(ud, ggd, ggl, delta)
    oubar, otau = OF.sfs(u, god, gol, delta)
    assert rel(ggl.download_spectral(ubar), oubar) < TOL64
    assert rel(ggl.download_spectral(tau), otau) < TOL64


# ---------------------------------------------------------------- closures
@pytest.mark.parametrize("D,n", [(3, 16), (2, 32)])
def test_closures_parity(D, n):
    import symcode as sc

    print_case("closures", D, n)
    go, gg = grids(D, n)
    u = oracle_solenoidal(go, 12, trunc=True)
    ud = gg.upload_spectral(u)
    delta = 2 * np.pi / n * 2
    GG = sc.gradient_phys(ud, gg)
    oGG = OC.gradient_phys(u, go)
    assert rel(gg.download_physical(GG), oGG) < TOL64
    tau = sc.create_clark(delta, gg)(ud, GG)
    assert rel(gg.download_physical(tau), OC.clark(oGG, delta, D)) < TOL64
    tau = sc.create_smagorinsky(0.1, delta, gg)(ud, GG)
    assert rel(gg.download_physical(tau), OC.smagorinsky(oGG, 0.1, delta, D)) < TOL64
    m = sc.create_dynamic_smagorinsky(delta, gg)
    tau = m(ud, GG)
    otau, oc = OC.dynamic_smagorinsky(u, delta, go, return_coeff=True)
    assert m.last_coefficient == pytest.approx(oc, rel=1e-11)
    assert rel(gg.download_physical(tau), otau) < 1e-11
    nu = 0.01
    for model, omodel in (
        (sc.create_clark(delta, gg), lambda u_, G_: OC.clark(G_, delta, D)),
        (sc.create_smagorinsky(0.1, delta, gg), lambda u_, G_: OC.smagorinsky(G_, 0.1, delta, D)),
        (sc.create_dynamic_smagorinsky(delta, gg), lambda u_, G_: OC.dynamic_smagorinsky(u_, delta, go)),
    ):
        du = sc.les_rhs(gg.vectorfield(), ud, gg, model, nu)
        assert rel(gg.download_spectral(du), OC.les_rhs(u, go, nu, omodel)) < 1e-11, type(model).__name__


@pytest.mark.parametrize("D,n", [(3, 16), (2, 32)])
def test_verstappen_bardina_parity(D, n):
    """create_verstappen (3D only) and create_bardina: tau and les! with
    them (closure ids 5 and 6) against the oracle, plus one LES RK4 step"""
    import symcode as sc

    print_case("verstappen bardina", D, n)
    go, gg = grids(D, n)
    u = oracle_solenoidal(go, 13, trunc=True)
    ud = gg.upload_spectral(u)
    delta = 2 * np.pi / n * 2
    GG = sc.gradient_phys(ud, gg)
    oGG = OC.gradient_phys(u, go)
    nu = 0.01
    models = [(sc.create_bardina(2.0, delta, gg), lambda u_, G_: OC.bardina(u_, 2.0, delta, go))]
    if D == 3:
        models.append((sc.create_verstappen(1.0, delta, gg), lambda u_, G_: OC.verstappen(G_, 1.0, delta)))
    else:
        with pytest.raises(ValueError):
            sc.create_verstappen(1.0, delta, gg)
    for model, omodel in models:
        name = type(model).__name__
        tau = model(ud, GG)
        assert rel(gg.download_physical(tau), omodel(u, oGG)) < TOL64, name
        du = sc.les_rhs(gg.vectorfield(), ud, gg, model, nu)
        assert rel(gg.download_spectral(du), OC.les_rhs(u, go, nu, omodel)) < 1e-11, name
        v = ud.clone()
        sc.les_rk4(v, 0.01, gg, model, nu)
        ref = S.rk4(lambda w: OC.les_rhs(w, go, nu, omodel), u, 0.01, go)
        assert rel(gg.download_spectral(v), ref) < 1e-11, name


# ----------------------------------------------------------------- FP32 path
@pytest.mark.parametrize("D,n", [(3, 16), (3, 30), (2, 810)])
def test_fp32_rhs(D, n):
    import symcode as sc

    print_case("fp32", D, n)
    go, gg = grids(D, n, prec=32)
    u = oracle_solenoidal(go, 13)
    du = sc.convectiondiffusion(gg.vectorfield(), gg.upload_spectral(u), gg, 0.02)
    assert rel(gg.download_spectral(du), S.convectiondiffusion(u, go, 0.02)) < TOL32
    x = rand_phys(go.phys_shape, 2)
    s = sc.to_spec(gg.spec(1), gg.upload_physical(x), gg)
    assert rel(gg.download_spectral(s)[0], S.to_spec(x, go)) < 1e-6


@pytest.mark.parametrize("n", [16, 30])
def test_fp32_rk4_3d(n):
    """the FP32 RK4 step in 3D (fused last pass and stage update) against the
    FP64 oracle, band forcing on"""
    import symcode as sc

    print_case("fp32 rk4", 3, n)
    go, gg = grids(3, n, prec=32)
    u = oracle_solenoidal(go, 15, trunc=True)
    sc.set_band_forcing(gg, 0.1, 2)
    nu, dt = 0.01, 0.02
    f = lambda v: S.convectiondiffusion(v, go, nu, forcing=(0.1, 2))
    ud = gg.upload_spectral(u)
    ref = u
    for k in range(3):
        ref = S.rk4(f, ref, dt, go)
        sc.rk4(ud, dt, gg, nu)
        assert rel(gg.download_spectral(ud), ref) < TOL32 * (k + 1), k
    sc.set_band_forcing(gg, 0.0)


@pytest.mark.parametrize("prec", [64, 32])
def test_rk4_810_fused_matches_unfused(prec):
    """810^3: the per-step RK4 entry point (fused five-pass right-hand side
    writing the divergence, then k_rk4_update_lin adding the viscous and
    forcing terms, accumulating and projecting) against the unfused path (the
    LES step without a closure: generic nonlinearity, divergence and viscous
    pass, separate projected update), band forcing on.  The oracle cannot run
    this size in test time; the two paths share no pass after the inverse
    transforms.  The multi-step path (k_z_fwd_rk4) is compared with this one
    in test_gpu_rk4_steps.py."""
    import gc

    import symcode as sc

    print_case("rk4 810 fused vs unfused", prec)
    dtype = torch.float64 if prec == 64 else torch.float32
    g = sc.Grid(3, 810, dtype=dtype)
    u = sc.randomfield(sc.peak_profile, g, seed=3, totalenergy=0.2, kpeak=4)
    sc.set_band_forcing(g, 0.1, 2)
    nu = 2.7e-4
    dt = 0.3 * sc.propose_timestep(u, g, nu)
    a = u.clone()
    sc.rk4(a, dt, g, nu)
    b = u.clone()
    sc.les_rk4(b, dt, g, None, nu)
    del u
    err = float(torch.linalg.norm((a - b).flatten()) / torch.linalg.norm(b.flatten()))
    print("rel diff", err, flush=True)
    assert err < (1e-12 if prec == 64 else TOL32)
    del a, b, g
    gc.collect()
    torch.cuda.empty_cache()


@pytest.mark.parametrize("D,prec", [(3, 64), (3, 32), (2, 32)])
def test_nonlinearity_810_against_separate_transforms(D, prec):
    """810^D: nonlinearity! through the fused x-pass (C2R, products and R2C
    of every row in one kernel, persistent with staged rows) against
    to_spec(v_i v_j) of v = to_phys(twothirds(u)) through the separate x
    transforms; and the CFL step's fused max |v| against propose_timestep."""
    import gc

    import symcode as sc

    print_case("nonlinearity 810 fused x-pass vs separate transforms", D, prec)
    g = sc.Grid(D, 810, dtype=torch.float64 if prec == 64 else torch.float32)
    u = sc.randomfield(sc.peak_profile, g, seed=5, totalenergy=0.2, kpeak=4)
    sig = sc.nonlinearity(g.tensorfield(), u, g)
    t = u.clone()
    sc.twothirds(t, g)
    v = sc.to_phys(g.spacevectorfield(), t, g)
    del t
    prod, spec = g.spacescalarfield(), g.scalarfield()
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0)] if D == 3 else [(0, 0), (1, 1), (0, 1)]
    for q, (i, j) in enumerate(pairs):
        torch.mul(v[i:i + 1], v[j:j + 1], out=prod)
        sc.to_spec(spec, prod, g)
        err = float(torch.linalg.norm((sig[q:q + 1] - spec).flatten()) / torch.linalg.norm(spec.flatten()))
        print("component", q, "rel diff", err, flush=True)
        assert err < (1e-13 if prec == 64 else 5e-6), q
    del sig, v, prod, spec
    nu, cfl = 2.7e-4, 0.3
    dto = cfl * sc.propose_timestep(u, g, nu)
    dtg = sc.rk4_cfl(u, g, nu, cfl)
    assert dtg == pytest.approx(dto, rel=1e-12 if prec == 64 else 1e-5)
    del u, g
    gc.collect()
    torch.cuda.empty_cache()


def test_config1_2d_256_full_length():
    """The reference's CPU-runnable configuration at full length: 2D 256^2
    FP64, E(k) ~ k^4 exp(-2 (k/5)^2) at unit energy (seed 0), nu = 1e-3,
    dt = 1e-3, no closure or forcing; 1000 RK4 steps and 300 Wray3 steps on
    the GPU against the oracle: the energy every 100 steps within 1e-12 and
    the final shell spectrum within 1e-11 relative"""
    import symcode as sc

    print_case("config 1: 2D 256 RK4 1000 steps, Wray3 300 steps")
    go, gg = grids(2, 256)
    prof = lambda k: S.peak_profile(k, 5)
    nu, dt = 1e-3, 1e-3
    f = lambda v: S.convectiondiffusion(v, go, nu)
    for ostep, gstep, nsteps in ((S.rk4, sc.rk4, 1000), (S.wray3, sc.wray3, 300)):
        ud = sc.randomfield(lambda k: prof(k), gg, seed=0, totalenergy=1.0)
        ref = S.randomfield(prof, go, seed=0, totalenergy=1.0)
        assert rel(gg.download_spectral(ud), ref) < TOL64
        for i in range(1, nsteps + 1):
            ref = ostep(f, ref, dt, go)
            gstep(ud, dt, gg, nu)
            if i % 100 == 0:
                assert sc.energy(ud, gg) == pytest.approx(S.energy(ref, go), rel=1e-12), i
        k, e = sc.spectrum(ud, gg)
        ko, eo = S.spectrum(ref, go)
        assert np.allclose(k, ko, rtol=1e-14)
        assert rel(e, eo) < 1e-11
        print(gstep.__name__, "energy", S.energy(ref, go), "spectrum rel", rel(e, eo), flush=True)


# ------------------------------------------------------- CFL-driven steppers
@pytest.mark.parametrize("D,n", [(3, 16), (3, 30), (2, 64)])
def test_cfl_steppers(D, n):
    """dt = min(cfl * propose_timestep(u), dtmax) computed on the device from
    the first stage, with band forcing inside the fused RHS."""
    import symcode as sc

    print_case("cfl", D, n)
    go, gg = grids(D, n)
    u = oracle_solenoidal(go, 14, trunc=True)
    nu, cfl = 0.01, 0.3
    sc.set_band_forcing(gg, 0.1, 2)
    f = lambda v: S.convectiondiffusion(v, go, nu, forcing=(0.1, 2))
    for ostep, gstep in ((S.rk4, sc.rk4_cfl), (S.wray3, sc.wray3_cfl)):
        ud = gg.upload_spectral(u)
        ref = u
        for k in range(3):
            dto = cfl * S.propose_timestep(ref, go, nu)
            dtg = gstep(ud, gg, nu, cfl)
            assert dtg == pytest.approx(dto, rel=1e-12)
            ref = ostep(f, ref, dto, go)
            assert rel(gg.download_spectral(ud), ref) < TOL64 * (k + 1)
    assert sc.rk4_cfl(ud, gg, nu, cfl, dtmax=1e-4) == 1e-4
    sc.set_band_forcing(gg, 0.0)


@pytest.mark.parametrize("kind", ["none", "clark", "smag", "dynsmag"])
def test_les_rollout(kind):
    """solve_les! (src/les.jl:188-247): CFL Wray3 with the closure, the shell
    clamp and save-time truncation, against the oracle's loop."""
    import symcode as sc

    print_case("les", kind)
    D, n = 3, 16
    go, gg = grids(D, n)
    u = oracle_solenoidal(go, 15, trunc=True)
    nu, cfl = 0.005, 0.35
    delta = 2 * np.pi / n * 2
    models = {
        "none": (None, None),
        "clark": (sc.create_clark(delta, gg), lambda u_, G: OC.clark(G, delta, D)),
        "smag": (sc.create_smagorinsky(0.1, delta, gg), lambda u_, G: OC.smagorinsky(G, 0.1, delta, D)),
        "dynsmag": (sc.create_dynamic_smagorinsky(delta, gg), lambda u_, G: OC.dynamic_smagorinsky(u_, delta, go)),
    }
    gm, om = models[kind]
    times = [0.0, 0.05, 0.1]
    ud = gg.upload_spectral(u)
    snaps = []
    sc.solve_les(ud, times, gg, nu, gm, cfl, forced=True, onsnapshot=lambda i, v: snaps.append(gg.download_spectral(v)))
    # oracle loop
    ref = u.copy()
    shells = S.energy_shells(ref, go)
    f = (lambda v: S.convectiondiffusion(v, go, nu)) if om is None else (lambda v: OC.les_rhs(v, go, nu, om))
    t = times[0]
    osnaps = []
    for i, tstop in enumerate(times):
        if i > 0:
            while t < tstop:
                dt = min(cfl * S.propose_timestep(ref, go, nu), tstop - t)
                t += dt
                ref = S.wray3(f, ref, dt, go)
                ref = S.maintain_shell_energy(ref, go, shells)
        ref = np.stack([S.twothirds(c, go) for c in ref])
        osnaps.append(ref)
    assert len(snaps) == len(osnaps)
    for a, b in zip(snaps, osnaps):
        assert rel(a, b) < 1e-10


# ------------------------------------------------------------------- adjoint
@pytest.mark.parametrize("D,n,forced", [(3, 16, False), (3, 16, True), (3, 30, False), (2, 64, False)])
def test_rhs_vjp_parity(D, n, forced):
    import symcode as sc
    from oracle import adjoint as OA

    print_case("rhs_vjp", D, n, forced)
    go, gg = grids(D, n)
    u = oracle_solenoidal(go, 16, trunc=True)
    w = oracle_solenoidal(go, 17)
    nu = 0.02
    forcing = (0.1, 2) if forced else None
    if forced:
        sc.set_band_forcing(gg, 0.1, 2)
    out = sc.rhs_vjp(gg.vectorfield(), gg.upload_spectral(u), gg.upload_spectral(w), gg, nu)
    assert rel(gg.download_spectral(out), OA.rhs_vjp(u, w, go, nu, forcing)) < TOL64
    sc.set_band_forcing(gg, 0.0)


@pytest.mark.parametrize("scheme", ["rk4", "wray3"])
def test_step_vjp_parity_and_dot_product(scheme):
    """GPU step adjoint vs the oracle adjoint, and a dot-product test against
    central differences of the GPU step itself."""
    import symcode as sc
    from oracle import adjoint as OA

    print_case("step_vjp", scheme)
    D, n = 3, 16
    go, gg = grids(D, n)
    u = oracle_solenoidal(go, 18, trunc=True)
    ub = oracle_solenoidal(go, 19)
    du = oracle_solenoidal(go, 20, trunc=True)
    nu, dt = 0.02, 0.05
    vjp = sc.rk4_vjp if scheme == "rk4" else sc.wray3_vjp
    ovjp = OA.rk4_vjp if scheme == "rk4" else OA.wray3_vjp
    step = sc.rk4 if scheme == "rk4" else sc.wray3
    got = gg.download_spectral(vjp(gg.upload_spectral(u), gg.upload_spectral(ub), dt, gg, nu))
    assert rel(got, ovjp(u, ub, dt, go, nu)) < 1e-11
    eps = 1e-4
    a = gg.upload_spectral(u + eps * du)
    b = gg.upload_spectral(u - eps * du)
    step(a, dt, gg, nu)
    step(b, dt, gg, nu)
    Jdu = (gg.download_spectral(a) - gg.download_spectral(b)) / (2 * eps)
    lhs = sum(S.spectraldot(ub[i], Jdu[i], go) for i in range(D)).real
    rhs_ = sum(S.spectraldot(got[i], du[i], go) for i in range(D)).real
    assert lhs == pytest.approx(rhs_, rel=1e-7)


def test_autograd_step_matches_finite_differences():
    """torch.autograd through symcode.autograd.rk4_step (two steps) against
    central differences along 4 random directions."""
    import symcode as sc

    print_case("autograd")
    D, n = 3, 16
    go, gg = grids(D, n)
    u = gg.upload_spectral(oracle_solenoidal(go, 21, trunc=True))
    target = gg.upload_spectral(oracle_solenoidal(go, 22, trunc=True))
    nu, dt = 0.02, 0.04

    def loss(v):
        for _ in range(2):
            v = sc.autograd.rk4_step(v, dt, gg, nu)
        return torch.sum(torch.abs(v - target) ** 2)

    x = u.clone().requires_grad_(True)
    L = loss(x)
    (grad,) = torch.autograd.grad(L, x)
    rng = np.random.default_rng(0)
    for _ in range(4):
        d = gg.upload_spectral(oracle_solenoidal(go, int(rng.integers(1000)), trunc=True))
        eps = 1e-4
        with torch.no_grad():
            fd = (loss(u + eps * d) - loss(u - eps * d)) / (2 * eps)
        an = torch.sum((torch.conj(grad) * d).real)
        assert float(an) == pytest.approx(float(fd), rel=1e-6)
