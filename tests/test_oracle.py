"""Pins for the CPU oracle: every oracle function is checked here against the
reference's own unit tests (KATs and property tests in test/*.jl of the
reference), against closed forms, or against brute force on tiny inputs."""
import itertools

import numpy as np
import pytest

from oracle import closures as OC
from oracle import filtering as OF
from oracle import rng as ORNG
from oracle import spectral as S


def rand_phys(g, seed, ncomp=None):
    r = np.random.default_rng(seed)
    shp = g.phys_shape if ncomp is None else (ncomp,) + g.phys_shape
    return r.standard_normal(shp)


def solenoidal(g, seed, trunc=True):
    u = np.stack([S.to_spec(rand_phys(g, seed + i), g) for i in range(g.D)])
    u = S.project(u, g)
    if trunc:
        u = np.stack([S.twothirds(u[i], g) for i in range(g.D)])
    return u


# ---------------------------------------------------------------- grid / FFT
def test_grid_ndrange():  # test/test_grid.jl
    g3 = S.Grid(3, 8, 1.0)
    g2 = S.Grid(2, 16)
    assert g3.spec_shape[::-1] == (5, 8, 8)
    assert g2.spec_shape[::-1] == (9, 16)
    assert g2.spacing == pytest.approx(2 * np.pi / 16)
    with pytest.raises(AssertionError):
        S.Grid(2, 7, 1.0)


def test_wavenumber_full():  # test/test_solver_kernels.jl:31-38
    g = S.Grid(2, 8, 4.0)
    kx = g.kfull[0][0, 1]
    ky = g.kfull[1][0, 0]
    assert kx == pytest.approx(2 * np.pi / 4.0) and ky == 0
    assert g.k2full()[0, 1] == pytest.approx(kx**2)


def test_fftfreq_nyquist_negative():
    k = S.fftfreq_int(8)
    assert list(k) == [0, 1, 2, 3, -4, -3, -2, -1]


@pytest.mark.parametrize("D,n", [(2, 6), (2, 8), (3, 4), (3, 6)])
def test_to_spec_bruteforce(D, n):
    """to_spec! = unnormalised R2C / n^D, checked against the explicit DFT sum."""
    g = S.Grid(D, n)
    x = rand_phys(g, 3)
    s = S.to_spec(x, g)
    idx = np.indices(g.phys_shape).reshape(D, -1)  # [z,y,x] order
    vals = x.reshape(-1)
    for kk in itertools.product(*[range(m) for m in g.spec_shape]):
        ph = sum(kk[a] * idx[a] for a in range(D))
        ref = np.sum(vals * np.exp(-2j * np.pi * ph / n)) / n**D
        assert abs(s[kk] - ref) < 1e-12


def test_constant_field_kat():  # test/test_fft.jl:21-23
    for g in (S.Grid(2, 16), S.Grid(3, 8, 1.0)):
        s = S.to_spec(np.full(g.phys_shape, 1.5), g)
        assert s.reshape(-1)[0] == pytest.approx(1.5)


def test_roundtrip_and_parseval():  # test/test_fft.jl:5-20
    for g in (S.Grid(2, 16), S.Grid(3, 8, 1.0)):
        x = rand_phys(g, 0)
        s = S.to_spec(x, g)
        assert np.allclose(S.to_phys(s, g), x, atol=1e-13)
        # Σ_w |û|² = ⟨|u|²⟩ (the solver's normalisation)
        assert S.getenergy(s, g) == pytest.approx(np.mean(x**2), rel=1e-13)


def test_dealias_idempotent():  # test/test_fft.jl:26-36
    g = S.Grid(3, 8)
    x = rand_phys(g, 1)
    once = S.dealias_phys(x, g)
    assert np.allclose(S.dealias_phys(once, g), once, atol=1e-13)


# ----------------------------------------------------------- spectral kernels
def _wavenumber_int_ref(g, I):
    """wavenumber_int restated with 1-based indices (src/solver.jl:167-172)."""
    def ff(i):
        return i - 1 - (0 if i <= (g.n + 1) >> 1 else g.n)
    return (I[0] - 1,) + tuple(ff(i) for i in I[1:])


@pytest.mark.parametrize("D,n", [(2, 24), (3, 12)])
def test_twothirds_bruteforce(D, n):  # test/test_solver_kernels.jl:6-17
    g = S.Grid(D, n)
    u = S.to_spec(rand_phys(g, 0), g)
    t = S.twothirds(u, g)
    kcut = n // 3
    for I in itertools.product(*[range(1, m + 1) for m in g.spec_shape[::-1]]):
        k = _wavenumber_int_ref(g, I)
        inband = any(k) and all(abs(c) <= kcut for c in k)
        val = t[tuple(i - 1 for i in I[::-1])]
        assert (val == u[tuple(i - 1 for i in I[::-1])]) if inband else val == 0


@pytest.mark.parametrize("D,n", [(2, 16), (3, 8)])
def test_project_divergence_free(D, n):  # test/test_solver_kernels.jl:19-28
    g = S.Grid(D, n)
    u = np.stack([S.to_spec(rand_phys(g, i), g) for i in range(D)])
    p = S.project(u, g)
    assert np.abs(S.vectordivergence(p, g)).max() < 1e-10
    assert np.allclose(S.project(p, g), p, atol=1e-14)
    assert p[(slice(None),) + (0,) * D].tolist() == u[(slice(None),) + (0,) * D].tolist()


def test_symmetrize_makes_hermitian():
    for g in (S.Grid(2, 8), S.Grid(3, 8)):
        u = S.random_spectral(g, 5, 1)[0]
        h = S.symmetrize(u, g)
        assert np.allclose(S.to_spec(S.to_phys(h, g), g), h, atol=1e-13)


# --------------------------------------------------------------------- RHS
@pytest.mark.parametrize("D,n", [(2, 32), (3, 16)])
def test_energy_budget(D, n):
    """Re Σ_w conj(u)·RHS(u) = -ε for truncated solenoidal u (alias-free when
    3 does not divide n): the nonlinear term conserves energy."""
    g = S.Grid(D, n)
    u = solenoidal(g, 7)
    nu = 3e-2
    du = S.convectiondiffusion(u, g, nu)
    lhs = sum(S.spectraldot(u[i], du[i], g) for i in range(D)).real
    eps = S.get_dissipation(u, nu, g)
    assert lhs == pytest.approx(-eps, rel=1e-12, abs=1e-14)
    # and the nonlinear part alone carries energy around without changing it
    dn = S.tensordivergence(S.nonlinearity(u, g), g)
    scale = np.sqrt(sum(S.getenergy(d, g) for d in dn))
    assert abs(sum(S.spectraldot(u[i], dn[i], g) for i in range(D)).real) < 1e-13 * scale


def _swap_yz(u):
    """y <-> z permutation of a 3D vector field [comp, z, y, x]."""
    return np.stack([u[0], u[2], u[1]]).transpose(0, 2, 1, 3)


def _reflect_z(u, n):
    idx = (-np.arange(n)) % n
    v = u[:, idx]
    v[2] = -v[2]
    return v


def _strip_nyquist(u, n):
    v = u.copy()
    v[:, n // 2] = 0
    v[:, :, n // 2] = 0
    v[..., n // 2] = 0
    return v


def test_rhs_equivariance_3d():
    """RHS commutes with the cube symmetries y<->z and z -> -z (Nyquist planes,
    where k = -n/2 is not reflection invariant, excluded)."""
    n = 12
    g = S.Grid(3, n)
    u = solenoidal(g, 11)
    f = lambda v: S.convectiondiffusion(v, g, 1e-2)
    a = _strip_nyquist(f(_swap_yz(u)), n)
    b = _strip_nyquist(_swap_yz(f(u)), n)
    assert np.abs(a - b).max() < 1e-13 * np.abs(b).max()
    a = _strip_nyquist(f(_reflect_z(u, n)), n)
    b = _strip_nyquist(_reflect_z(f(u), n), n)
    assert np.abs(a - b).max() < 1e-13 * np.abs(b).max()


def test_rhs_translation():
    """A shift by one grid cell along x is a phase ramp; the RHS commutes."""
    n = 12
    g = S.Grid(3, n)
    u = solenoidal(g, 12)
    ph = np.exp(-1j * g.kint[0] * 2 * np.pi / n)
    f = lambda v: S.convectiondiffusion(v, g, 1e-2)
    a = f(u * ph)
    b = f(u) * ph
    assert np.abs(a - b).max() < 1e-13 * np.abs(b).max()


def test_tgv_stability_polynomials():
    """2D Taylor-Green: the nonlinear term is a gradient, so every step scales
    the amplitude by the stability polynomial R(-2νΔt) exactly."""
    g = S.Grid(2, 16)
    u = S.taylorgreen(g)
    nu, dt, m = 1e-3, 1e-3, 25
    z = -2 * nu * dt
    f = lambda v: S.convectiondiffusion(v, g, nu)
    for step, R in ((S.wray3, 1 + z + z**2 / 2 + z**3 / 6),
                    (S.rk4, 1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24)):
        v = u
        for _ in range(m):
            v = step(f, v, dt, g)
        assert np.allclose(v, u * R**m, atol=1e-15, rtol=1e-13)
    # closed form at 1000 Wray3 steps (BASELINE config-1 settings)
    assert (1 + z + z**2 / 2 + z**3 / 6) ** 1000 == pytest.approx(0.99800199866734, rel=1e-13)


def test_shear_wave_3d():
    """u = (0, 0, sin(k x)) has no nonlinear term and decays by R(-ν k² Δt)."""
    n, kk = 16, 3
    g = S.Grid(3, n)
    x = (np.arange(n) + 0.5) * 2 * np.pi / n
    uz = np.broadcast_to(np.sin(kk * x)[None, None, :], g.phys_shape)
    u = np.stack([np.zeros(g.spec_shape, complex), np.zeros(g.spec_shape, complex), S.to_spec(uz, g)])
    nu, dt = 0.05, 0.01
    du = S.convectiondiffusion(u, g, nu)
    assert np.allclose(du, -nu * kk**2 * u, atol=1e-14)
    v = S.rk4(lambda w: S.convectiondiffusion(w, g, nu), u, dt, g)
    z = -nu * kk**2 * dt
    assert np.allclose(v, u * (1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24), atol=1e-15)


def test_propose_timestep_closed_form():
    g = S.Grid(2, 16)
    u = S.taylorgreen(g, V0=2.0)
    nu = 1e-2
    h = g.l / g.n
    vmax = np.sqrt(sum(S.to_phys(S.twothirds(u[i], g), g) ** 2 for i in range(2))).max()
    x = (np.arange(16) + 0.5) * h
    X, Y = x[None, :], x[:, None]
    exact = 2.0 * np.sqrt(np.sin(X) ** 2 * np.cos(Y) ** 2 + np.cos(X) ** 2 * np.sin(Y) ** 2).max()
    assert vmax == pytest.approx(exact, rel=1e-12)
    assert S.propose_timestep(u, g, nu) == pytest.approx(min(h / vmax, h**2 / 2 / 2 / nu))


def test_band_forcing_injection():
    g = S.Grid(3, 16)
    u = solenoidal(g, 3)
    f = S.band_forcing(u, g, 0.1, 2)
    inj = sum(S.spectraldot(u[i], f[i], g) for i in range(3)).real
    assert inj == pytest.approx(0.1, rel=1e-13)
    k2 = g.k2int()
    assert np.all(f[:, (k2 == 0) | (k2 > 4)] == 0)


# ------------------------------------------------------------- reductions
def test_spectrum_single_mode():
    g = S.Grid(3, 16)
    u = np.zeros((3,) + g.spec_shape, complex)
    u[1, 0, 0, 3] = 0.5  # kx = 3, counted twice (RFFT doubling)
    u[0, 4, 0, 0] = 0.25  # kz = 4 on the kx = 0 plane, counted once
    k, s = S.spectrum(u, g)
    assert s[2] == pytest.approx(0.25) and s[3] == pytest.approx(0.25 ** 2 / 2)
    assert np.count_nonzero(s) == 2
    assert k[0] == pytest.approx(1.0)


def test_maintain_shell_energy_restores():
    g = S.Grid(3, 12)
    u = solenoidal(g, 4)
    shells = S.energy_shells(u, g)
    v = S.maintain_shell_energy(1.7 * u, g, shells)
    for (i, e), (_, e2) in zip(shells, S.energy_shells(v, g)):
        assert e2 == pytest.approx(e, rel=1e-13)
    sh = S.shell_of_modes(g)
    assert np.allclose(v[:, sh > 2], 1.7 * u[:, sh > 2])


def test_randomfield_spectrum():
    g = S.Grid(3, 16)
    prof = lambda k: S.peak_profile(k, 4)
    u = S.randomfield(prof, g, seed=1, totalenergy=1.0)
    assert np.abs(S.vectordivergence(u, g)).max() < 1e-12
    kdiag = int(np.floor(np.sqrt(3) * 8))
    tot = sum(prof(k) for k in range(kdiag + 1))
    e = S.shell_energies_all = None
    sh = S.shell_of_modes(g)
    w = g.rfft_weight()
    E = w * np.sum(np.abs(u) ** 2, axis=0) / 2
    for k in range(1, 5):  # shells fully inside the 2/3 box
        assert E[sh == k].sum() == pytest.approx(prof(k) / tot, rel=1e-12)
    # Hermitian on the kx = 0 plane: R2C(C2R(u)) returns u
    back = np.stack([S.to_spec(S.to_phys(u[i], g), g) for i in range(3)])
    assert np.allclose(back, u, atol=1e-14)


def test_rng_stream():
    a = ORNG.complex_normals(3, 2, (40, 40))
    b = ORNG.complex_normals(3, 2, (40, 40))
    assert np.array_equal(a, b)
    assert abs(a.real.var() - 0.5) < 0.05 and abs(a.imag.var() - 0.5) < 0.05
    assert not np.array_equal(a[0], a[1])
    # first value of the stream (fixes the splitmix64 / Box-Muller recipe)
    key = ORNG.mix64(np.uint64(3) * ORNG.G + np.uint64(1))
    assert int(key) == int(ORNG.mix64(np.uint64((3 * 0x9E3779B97F4A7C15 + 1) % 2**64)))


# ---------------------------------------------------------------- filters
def test_gaussian_closed_form():
    g = S.Grid(3, 16, 3.0)
    u = S.to_spec(rand_phys(g, 2), g)
    d = 0.4
    v = OF.gaussianfilter(u, d, g)
    k2 = g.k2full()
    ref = np.where(g.k2int() < 9, 1.0, np.exp(-(d**2) * k2 / 24)) * u
    assert np.allclose(v, ref, rtol=1e-15, atol=0)
    assert np.array_equal(v[g.k2int() < 9], u[g.k2int() < 9])


def test_cutoff_index_bruteforce():
    """cutoff_index restated with the reference's 1-based arithmetic."""
    nbar, n = 8, 24
    for is1 in (True, False):
        for i1 in range(1, (nbar // 2 + 1 if is1 else nbar) + 1):
            imax = nbar // 2 + (1 if is1 else 0)
            ref1 = n - nbar + i1 if i1 > imax else i1
            assert OF.cutoff_index(nbar, n, i1 - 1, is1) == ref1 - 1
    gd, gl = S.Grid(3, 24), S.Grid(3, 8)
    u = S.to_spec(rand_phys(gd, 9), gd)
    ub = OF.cutoff(u, gl, gd)
    # same wavenumber on both grids
    for I in itertools.product(range(8), range(8), range(5)):
        kc = (S.fftfreq_int(8)[I[0]], S.fftfreq_int(8)[I[1]], I[2])
        J = (kc[0] % 24, kc[1] % 24, kc[2])
        assert ub[I] == u[J]


def test_sfs_identity_grid():
    """With the LES grid equal to the DNS grid, no filter, and a truncated u,
    the SFS stress vanishes."""
    g = S.Grid(3, 16)
    u = solenoidal(g, 5)
    ubar, tau = OF.sfs(u, g, g, None)
    assert np.allclose(ubar, u)
    assert np.abs(tau).max() < 1e-15
    gl = S.Grid(3, 8)
    ubar, tau = OF.sfs(u, g, gl, 2 * np.pi / 8 * 2)
    tr = tau[0] + tau[1] + tau[2]
    assert np.abs(tr).max() < 1e-15
    assert np.all(S.twothirds(tau[0], gl) == tau[0])


# --------------------------------------------------------------- closures
def _const_grad(D, n, entries):
    GG = np.zeros((D * D,) + (n,) * D)
    for k, v in entries.items():
        GG[k] = v
    return GG


def test_clark_kat():  # test/test_closures.jl:4-17
    # 2D nonsym order xx, yx, xy, yy; G.xy = 2
    tau = OC.clark(_const_grad(2, 4, {2: 2.0}), 0.5, 2)
    assert np.allclose(tau[0], 0.5**2 / 12 * 4) and np.allclose(tau[1:], 0)
    # 3D analogue: G.xz = 3 (index 6) -> tau_xx = Δ²/12 · 9
    tau = OC.clark(_const_grad(3, 4, {6: 3.0}), 0.5, 3)
    assert np.allclose(tau[0], 0.5**2 / 12 * 9) and np.allclose(tau[1:], 0)


def test_smagorinsky_kat():  # test/test_closures.jl:19-33
    D, a, cs, d = 2, 1.7, 0.17, 0.3
    tau = OC.smagorinsky(_const_grad(2, 4, {2: a}), cs, d, D)
    assert np.allclose(tau[:2], 0)
    assert np.allclose(tau[2], -2 * cs**2 * d**2 * abs(a) * (a / 2))
    # 3D pure shear G.yz = a (index 7): tau_yz = -2 C_S²Δ²|a|(a/2)
    tau = OC.smagorinsky(_const_grad(3, 4, {7: a}), cs, d, 3)
    assert np.allclose(tau[4], -2 * cs**2 * d**2 * abs(a) * (a / 2))
    assert np.allclose(np.delete(tau, 4, axis=0), 0)


def test_dynsmag_zero_mm():  # test/test_closures.jl:35-53
    M = np.zeros((3, 4, 4))
    L = np.ones((3, 4, 4))
    ml, mm = OC.ml_mm(M, L, 2)
    assert np.all(mm == 0)
    c = OC.dynamic_coefficient(float(ml.sum()), float(mm.sum()))
    assert c == 0 and not np.isnan(c)


def test_dynsmag_coefficient_sign():
    """c = max(-<ML>/<MM>, 0) is non-negative and scales tau = c·m1."""
    g = S.Grid(3, 16)
    u = solenoidal(g, 8)
    tau, c = OC.dynamic_smagorinsky(u, 2 * np.pi / 16, g, return_coeff=True)
    assert c >= 0
    assert np.isfinite(tau).all()


def test_les_rhs_no_closure_is_dns_rhs():
    g = S.Grid(3, 12)
    u = solenoidal(g, 2)
    zero = lambda u, GG: np.zeros((6,) + g.phys_shape)
    assert np.allclose(OC.les_rhs(u, g, 1e-2, zero), S.convectiondiffusion(u, g, 1e-2), atol=1e-15)


def test_verstappen_closed_form():
    """Q-R model on constant gradients: a trace-free diagonal strain gives
    q = (a² + b² + (a+b)²)/2, r = -ab(a+b); pure shear has r = 0 (tr S³ = 0)."""
    a, b, c, d = 0.7, -1.9, 1.0, 0.3
    tau = OC.verstappen(_const_grad(3, 4, {0: a, 4: b, 8: -a - b}), c, d)
    q = (a * a + b * b + (a + b) ** 2) / 2
    nu = c**2 * d**2 * abs(a * b * (a + b)) / q
    assert np.allclose(tau[0], -2 * nu * a, rtol=1e-14)
    assert np.allclose(tau[1], -2 * nu * b, rtol=1e-14)
    assert np.allclose(tau[2], 2 * nu * (a + b), rtol=1e-14)
    assert np.allclose(tau[3:], 0)
    assert np.allclose(OC.verstappen(_const_grad(3, 4, {3: 1.3}), c, d), 0, atol=1e-16)


@pytest.mark.parametrize("D", [2, 3])
def test_bardina_single_mode(D):
    """u_x = cos(3 x_last): before the trace removal tau_xx =
    -g3²/2 + (g6 - g3²) cos(6 x_last)/2 (g_k the filter's transfer at |k|;
    the filtered u u loses its mean to twothirds!, which also zeroes k = 0,
    while (G~ u)(G~ u) keeps g3²/2), all other components zero"""
    n, s, d = 32, 2.0, 0.3
    g = S.Grid(D, n)
    x = np.arange(n) * (2 * np.pi / n)
    z = x.reshape((n,) + (1,) * (D - 1))
    ux = np.broadcast_to(np.cos(3 * z), g.phys_shape).copy()
    u = np.stack([S.to_spec(ux, g)] + [S.to_spec(np.zeros(g.phys_shape), g) for _ in range(D - 1)])
    tau = OC.bardina(u, s, d, g)
    g3, g6 = np.exp(-(s * d) ** 2 * 9 / 24), np.exp(-(s * d) ** 2 * 36 / 24)
    T = -0.5 * g3**2 + 0.5 * (g6 - g3**2) * np.cos(6 * z)
    T = np.broadcast_to(T, g.phys_shape)
    assert np.allclose(tau[0], (D - 1) / D * T, atol=1e-14)
    assert np.allclose(tau[1], -T / D, atol=1e-14)
    if D == 3:
        assert np.allclose(tau[2], -T / 3, atol=1e-14)
        assert np.allclose(tau[3:], 0, atol=1e-14)
    else:
        assert np.allclose(tau[2], 0, atol=1e-14)


def test_eddy_viscosity_dissipation_and_transfer():
    """tau = -nu_t (G + G^T): eps = 2 nu_t S:S pointwise, and for solenoidal u
    the transfer clo = -nu_t k^2 u gives -nu_t sum_shell w k^2 |u|^2"""
    from oracle import analysis as OA

    g = S.Grid(3, 16)
    u = solenoidal(g, 31)
    u = np.stack([S.twothirds(c, g) for c in u])
    nut = 0.037
    GG = OC.gradient_phys(u, g)
    G = OC.grad_matrix(GG, 3)
    Sm = (G + np.swapaxes(G, 0, 1)) / 2
    tau = OC.pack_sym(-2 * nut * Sm, 3)
    eps = OA.sfs_dissipation(tau, GG, 3)
    assert np.allclose(eps, 2 * nut * np.sum(Sm**2, axis=(0, 1)), rtol=1e-12, atol=1e-14)
    T = OA.shell_transfer(u, tau, g)
    e = g.rfft_weight() * g.k2full() * sum(np.abs(u[i]) ** 2 for i in range(3))
    sh = S.shell_of_modes(g)
    ref = np.array([-nut * e[sh == i].sum() for i in range(1, len(T) + 1)])
    assert np.allclose(T, ref, rtol=1e-11, atol=1e-14 * np.abs(ref).max())


def test_moments_closed_form():
    from oracle import analysis as OA

    m = OA.moments([1.0, 2.0, 3.0, 4.0])
    assert m["mean"] == 2.5 and m["median"] == 2.5 and m["backscatter"] == 0.0
    assert abs(m["std"] - np.sqrt(5 / 3)) < 1e-15 and abs(m["skewness"]) < 1e-15
    m = OA.moments([-1.0, 0.0, 0.0, 0.0, 5.0])
    s = np.sqrt(22.8 / 4)
    assert abs(m["mean"] - 0.8) < 1e-15 and m["median"] == 0.0 and m["backscatter"] == 0.2
    assert abs(m["std"] - s) < 1e-14
    assert abs(m["skewness"] - (-1.8**3 - 3 * 0.8**3 + 4.2**3) / 5 / s**3) < 1e-13
    assert OA.moments([2.0, 2.0, 2.0])["skewness"] == 0.0
