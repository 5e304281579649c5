"""Closed-form pins of the oracle's dynamic Smagorinsky model, filtered SFS
stress and filter Reynolds number, each with the plausible mistakes planted
in the oracle (monkeypatched, CPU only) to show that the pin catches them.

Fields are single-wavevector waves u = Re(a e^{iκz}), where every spectral
step (2/3 mask, Gaussian filter, cutoff) is a scalar per band, so L, M, m1,
m2, c and the SFS stress are explicit trigonometric expressions.

Dynamic Smagorinsky (src/closures.jl:188-271).  With a_x² + a_y² + 2a_z² = 0
the strain's bilinear square Ŝ:Ŝ vanishes, so S:S is constant and
|S| = √(2 S:S) = s0 = κ √(|a_z|² + (|a_x|² + |a_y|²)/2): the Smagorinsky
tensors are linear in S.  With a_z ≠ 0 the field is not solenoidal, so
tr S = ∂_z u_z ≠ 0 and the trace removal of L matters.  At κ = n/3 the
product band 2κ aliases onto -κ on the grid (the 2/3 mask keeps κ = ⌊n/3⌋),
so L has a band-κ part and <M:L> ≠ 0:
    L_ij = (g - g²)/2 Re(conj(a_i a_j) e^{iκz}) - g² Re(a_i conj(a_j))/2
    M    = μ S,  μ = -2 s0 (Δ² g - Δ̂² g²),  Δ̂² = (2Δ)² + Δ²
    <M:L_dev> = μ (g - g²)/4 · (2κ/3) Im(a_z³),   <M:M> = μ² s0²/2
    c = max(-<M:L_dev>/<M:M>, 0),   τ = c m1,   m1 = -2 Δ² s0 S
with g = exp(-(2Δ)² κ²/24) the test filter's transfer at κ.
"""
import numpy as np
import pytest

from oracle import closures as OC
from oracle import filtering as OF
from oracle import spectral as S


def wave(g, a, kappa):
    """u_i = Re(a_i e^{iκz}) sampled at z_m = 2πm/n (z is the first axis)."""
    n = g.n
    z = (np.arange(n) * (g.l / n)).reshape((n,) + (1,) * (g.D - 1))
    kz = 2 * np.pi / g.l * kappa
    phys = [np.broadcast_to(np.real(ai * np.exp(1j * kz * z)), g.phys_shape).copy() for ai in a]
    return np.stack([S.to_spec(p, g) for p in phys]), z


# -------------------------------------------------------- dynamic Smagorinsky
N_DS, KAPPA, DELTA = 12, 4, 0.3
A_Z = 0.8 * np.exp(-1j * np.pi / 6)
A_X = 1.0 + 0.0j
A_Y = np.sqrt(-2 * A_Z**2 - A_X**2 + 0j)


def dynsmag_closed_form():
    a = np.array([A_X, A_Y, A_Z])
    k = float(KAPPA)
    g = np.exp(-((2 * DELTA) ** 2) * k**2 / 24)
    dhat2 = (2 * DELTA) ** 2 + DELTA**2
    s0 = k * np.sqrt(abs(A_Z) ** 2 + (abs(A_X) ** 2 + abs(A_Y) ** 2) / 2)
    mu = -2 * s0 * (DELTA**2 * g - dhat2 * g**2)
    ml = mu * (g - g * g) / 4 * (2 * k / 3) * np.imag(A_Z**3)
    mm = mu**2 * s0**2 / 2
    c = max(-ml / mm, 0.0)
    return dict(a=a, g=g, s0=s0, mu=mu, c=c, dhat2=dhat2)


def _check_dynsmag(tol=1e-12):
    """Largest relative deviation of the oracle's L, M, m1, c and τ from the
    closed forms (0 to round-off for a correct oracle)."""
    gr = S.Grid(3, N_DS)
    cf = dynsmag_closed_form()
    a, gk, mu, s0 = cf["a"], cf["g"], cf["mu"], cf["s0"]
    u, z = wave(gr, a, KAPPA)
    tau, parts = OC.dynamic_smagorinsky(u, DELTA, gr, return_parts=True)
    e = np.exp(1j * KAPPA * z)
    Sf = {  # strain, packed xx, yy, zz, xy, yz, zx
        2: np.real(1j * KAPPA * a[2] * e),
        4: np.real(1j * KAPPA * a[1] / 2 * e),
        5: np.real(1j * KAPPA * a[0] / 2 * e),
    }
    Sp = np.stack([np.broadcast_to(Sf.get(i, np.zeros_like(z)), gr.phys_shape) for i in range(6)])
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0)]
    Lp = np.stack([
        np.broadcast_to((gk - gk * gk) / 2 * np.real(np.conj(a[i] * a[j]) * e)
                        - gk * gk * np.real(a[i] * np.conj(a[j])) / 2, gr.phys_shape)
        for i, j in pairs
    ])
    dev = {}
    sc = np.abs(Sp).max()
    dev["L"] = np.abs(parts["L"] - Lp).max() / np.abs(Lp).max()
    dev["M"] = np.abs(parts["M"] - mu * Sp).max() / (abs(mu) * sc)
    dev["m1"] = np.abs(parts["m1"] + 2 * DELTA**2 * s0 * Sp).max() / (2 * DELTA**2 * s0 * sc)
    dev["c"] = abs(parts["c"] - cf["c"]) / cf["c"]
    dev["tau"] = np.abs(tau + cf["c"] * 2 * DELTA**2 * s0 * Sp).max() / (cf["c"] * 2 * DELTA**2 * s0 * sc)
    return dev


def test_dynsmag_closed_form_field():
    cf = dynsmag_closed_form()
    # the construction: constant |S|, a non-solenoidal field, c > 0
    assert abs(A_X**2 + A_Y**2 + 2 * A_Z**2) < 1e-15
    assert cf["c"] > 1e-4 and cf["mu"] > 0
    dev = _check_dynsmag()
    assert max(dev.values()) < 1e-12, dev


def test_dynsmag_mutant_test_filter_width(monkeypatch):
    """Δ̂ = 2Δ instead of √(Δ̃² + Δ²) for m2."""
    orig = OC.smagorinsky_tensor
    calls = []

    def mutant(Sv, delta, D):
        calls.append(delta)
        return orig(Sv, 2 * calls[0] if len(calls) == 2 else delta, D)

    monkeypatch.setattr(OC, "smagorinsky_tensor", mutant)
    dev = _check_dynsmag()
    assert dev["M"] > 1e-2 and dev["c"] > 1e-2


def test_dynsmag_mutant_sign_of_m(monkeypatch):
    """M = m2 - G̃ m1 instead of G̃ m1 - m2: the clipped coefficient drops to 0."""
    tf, st = OC.test_filter_phys, OC.smagorinsky_tensor
    calls = []

    def neg_tensor(Sv, delta, D):
        calls.append(delta)
        r = st(Sv, delta, D)
        return -r if len(calls) % 2 == 0 else r  # the m2 call

    monkeypatch.setattr(OC, "test_filter_phys", lambda m, d, g: -tf(m, d, g))
    monkeypatch.setattr(OC, "smagorinsky_tensor", neg_tensor)
    _, parts = OC.dynamic_smagorinsky(wave(S.Grid(3, N_DS), dynsmag_closed_form()["a"], KAPPA)[0], DELTA,
                                      S.Grid(3, N_DS), return_parts=True)
    assert parts["c"] == 0.0
    dev = _check_dynsmag()
    assert dev["M"] > 1.0 and dev["c"] >= 1.0


def test_dynsmag_mutant_nondeviatoric_l(monkeypatch):
    """<M:L> instead of <M:L_dev> (M carries a trace here: tr S ≠ 0)."""

    def mutant(M, L, D):
        ml = M[0] * L[0] + M[1] * L[1] + M[2] * L[2] + 2 * (M[3] * L[3] + M[4] * L[4] + M[5] * L[5])
        mm = M[0] * M[0] + M[1] * M[1] + M[2] * M[2] + 2 * (M[3] * M[3] + M[4] * M[4] + M[5] * M[5])
        return ml, mm

    monkeypatch.setattr(OC, "ml_mm", mutant)
    dev = _check_dynsmag()
    assert abs(dev["c"] - 0.5) < 1e-9  # 3/2 of the coefficient


def test_dynsmag_m1_dealiased():
    """τ = c dealias(m1): on a broadband field |S|S has content beyond the 2/3
    box and at k = 0, τ has none; without the dealias it would."""
    g = S.Grid(3, 16)
    r = np.random.default_rng(3)
    u = np.stack([S.twothirds(S.to_spec(r.standard_normal(g.phys_shape), g), g) for _ in range(3)])
    u = S.project(u, g)
    tau, parts = OC.dynamic_smagorinsky(u, 2 * np.pi / 16, g, return_parts=True)
    assert parts["c"] > 0
    for t in tau:
        th = S.to_spec(t, g)
        assert np.abs(th - S.twothirds(th, g)).max() < 1e-14 * np.abs(th).max()
    # the undealiased m1 does have content outside the box
    Sh = S.strainrate(u, g)
    Sp = np.stack([S.to_phys(S.twothirds(s, g), g) for s in Sh])
    raw = OC.smagorinsky_tensor(Sp, 2 * np.pi / 16, 3)
    th = S.to_spec(raw[3], g)
    assert np.abs(th - S.twothirds(th, g)).max() > 1e-3 * np.abs(th).max()


def test_dynsmag_mutant_no_dealias(monkeypatch):
    monkeypatch.setattr(OC, "dealias_phys", lambda m, g: m)
    g = S.Grid(3, 16)
    r = np.random.default_rng(3)
    u = np.stack([S.twothirds(S.to_spec(r.standard_normal(g.phys_shape), g), g) for _ in range(3)])
    u = S.project(u, g)
    tau = OC.dynamic_smagorinsky(u, 2 * np.pi / 16, g)
    th = S.to_spec(tau[3], g)
    assert np.abs(th - S.twothirds(th, g)).max() > 1e-3 * np.abs(th).max()


def test_ml_mm_kat():
    """smagorinsky_ml_mm! at one point: M = diag(1, 0, 0) + M_xy = 0.5,
    L = diag(0, 0, 3) + L_yz = 2: L_dev = diag(-1, -1, 2), M:L_dev = -1
    (the yz entry meets M_yz = 0), M:M = 1 + 2·0.25."""
    M = np.array([1.0, 0, 0, 0.5, 0, 0]).reshape(6, 1)
    L = np.array([0, 0, 3.0, 0, 2.0, 0]).reshape(6, 1)
    ml, mm = OC.ml_mm(M, L, 3)
    assert ml[0] == -1.0 and mm[0] == 1.5
    ml, mm = OC.ml_mm(np.array([1.0, 0, 0.5]).reshape(3, 1), np.array([0, 2.0, 1.0]).reshape(3, 1), 2)
    assert ml[0] == -1.0 + 1.0 and mm[0] == 1.5  # L_dev = (-1, 1); 2·0.5·1
    assert OC.dynamic_coefficient(-2.0, 4.0) == 0.5 and OC.dynamic_coefficient(2.0, 4.0) == 0.0


# ------------------------------------------------------------ filtered sfs!
N_DNS, N_LES, KS, DS = 24, 18, 3, 0.5
A_S = np.array([1.0, 0.5j, 0.3 + 0.2j])


def _check_sfs():
    gd, gl = S.Grid(3, N_DNS), S.Grid(3, N_LES)
    u, _ = wave(gd, A_S, KS)
    ubar, tau = OF.sfs(u, gd, gl, DS)
    _, z = wave(gl, A_S, KS)
    g1 = np.exp(-DS**2 * KS**2 / 24)
    g2 = np.exp(-DS**2 * (2 * KS) ** 2 / 24)
    ub_ref, _ = wave(gl, g1 * A_S, KS)
    e2 = np.exp(2j * KS * z)
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0)]
    T = np.stack([np.broadcast_to((g2 - g1 * g1) / 2 * np.real(A_S[i] * A_S[j] * e2), gl.phys_shape)
                  for i, j in pairs])
    T = OF.make_tracefree(T, 3)
    tau_phys = np.stack([S.to_phys(t, gl) for t in tau])
    return (np.abs(ubar - ub_ref).max() / np.abs(ub_ref).max(),
            np.abs(tau_phys - T).max() / np.abs(T).max())


def test_sfs_filtered_closed_form():
    """σ̄₁ = G(cutoff(σ)) keeps the mean and scales band 2κ by g(2κ); σ̄₂ = ū ū
    = g(κ)² u u; the 2/3 mask removes the mean, so τ = tracefree((g(2κ) -
    g(κ)²)/2 Re(a_i a_j e^{2iκz})) and ū = g(κ) u (κ = 3 is filtered:
    |k|² = 9 is not below the forced-shell bound)."""
    du, dt = _check_sfs()
    assert du < 1e-14 and dt < 1e-12, (du, dt)


def test_sfs_mutant_unfiltered_sigma1(monkeypatch):
    orig = OF.gaussianfilter
    calls = []

    def mutant(u, delta, g):
        calls.append(1)
        return orig(u, delta, g) if len(calls) <= 3 else u  # σ̄₁ left unfiltered

    monkeypatch.setattr(OF, "gaussianfilter", mutant)
    du, dt = _check_sfs()
    assert du < 1e-14 and dt > 0.1


# ------------------------------------------------------- filter_reynolds
def test_filter_reynolds_closed_form():
    """ū = Re(a e^{iκz}) (+ a mode beyond the 2/3 box, which the reference's
    mask drops): Σ_ij <(∂_j ū_i)²> = κ²|a|²/2, so Re_Δ = Δ² κ |a| / (√2 ν)."""
    g = S.Grid(3, 16)
    a = np.array([0.3, -0.7j, 0.2 + 0.1j])
    u, _ = wave(g, a, 2)
    u2, _ = wave(g, np.array([0.5, 0.5, 0.5]), 7)  # 7 > ⌊16/3⌋
    visc, d = 2e-3, 0.4
    ref = d**2 * 2 * np.linalg.norm(a) / np.sqrt(2) / visc
    assert abs(OF.filter_reynolds(u, g, visc, d) - ref) < 1e-13 * ref
    assert abs(OF.filter_reynolds(u + u2, g, visc, d) - ref) < 1e-13 * ref
    # without the mask the extra mode would count
    G, _ = S.vectorgradient(u + u2, g)
    unmasked = d**2 * np.sqrt(sum(OF.getenergy(x, g) for x in G)) / visc
    assert unmasked > 1.5 * ref
