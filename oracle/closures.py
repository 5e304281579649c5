"""Oracle: classical LES closures and the LES right-hand side (restates
src/closures.jl and src/les.jl).  TEST INFRASTRUCTURE ONLY -- see
``oracle/__init__.py``.

Physical tensor fields are numpy arrays ``[comp, (z,) y, x]``.  The velocity
gradient ``GG`` uses the reference's non-symmetric order
``[xx, yx, zx, xy, yy, zy, xz, yz, zz]`` (tensorfield_nonsym) with
``G.ij = ∂u_i/∂x_j``; packed closure outputs use ``[xx, yy, zz, xy, yz, zx]``.
"""
import numpy as np

from .filtering import gaussianfilter, make_tracefree
from .spectral import (
    Grid,
    dealias_phys,
    nonlinearity,
    stress,
    strainrate,
    tensordivergence,
    test_filter_phys,
    to_phys,
    to_spec,
    twothirds,
    vectorgradient,
)


def grad_matrix(GG, D):
    """tensorat (src/closures.jl:6-11): G[i][j] = GG.ij as a (D, D, ...) array."""
    names = "xyz"[:D]
    order = ["xx", "yx", "xy", "yy"] if D == 2 else ["xx", "yx", "zx", "xy", "yy", "zy", "xz", "yz", "zz"]
    pos = {o: k for k, o in enumerate(order)}
    return np.stack([np.stack([GG[pos[names[i] + names[j]]] for j in range(D)]) for i in range(D)])


def pack_sym(tau, D):
    """store_symtensor! (src/closures.jl:14-29): upper triangle, packed."""
    if D == 2:
        return np.stack([tau[0, 0], tau[1, 1], tau[0, 1]])
    return np.stack([tau[0, 0], tau[1, 1], tau[2, 2], tau[0, 1], tau[1, 2], tau[2, 0]])


def clark(GG, delta, D):
    """clark_kernel! (src/closures.jl:37-42): τ = Δ²/12 · G Gᵀ."""
    G = grad_matrix(GG, D)
    tau = delta**2 / 12 * np.einsum("ik...,jk...->ij...", G, G)
    return pack_sym(tau, D)


def smagorinsky(GG, cs, delta, D):
    """smagorinsky_kernel! (src/closures.jl:44-51): τ = -2 C_S²Δ²√(2S:S) S."""
    G = grad_matrix(GG, D)
    S = (G + np.swapaxes(G, 0, 1)) / 2
    nu = cs**2 * delta**2 * np.sqrt(2 * np.sum(S**2, axis=(0, 1)))
    return pack_sym(-2 * nu * S, D)


def verstappen(GG, c, delta):
    """verstappen_kernel! (src/closures.jl:53-62), 3D: q = tr(S S)/2,
    r = tr(S S S)/3, ν = C²Δ²|r|/q, τ = -2νS."""
    G = grad_matrix(GG, 3)
    S = (G + np.swapaxes(G, 0, 1)) / 2
    SS = np.einsum("ik...,kj...->ij...", S, S)
    q = np.einsum("ii...->...", SS) / 2
    r = np.einsum("ik...,ki...->...", SS, S) / 3
    nu = c**2 * delta**2 * np.abs(r) / q
    return pack_sym(-2 * nu * S, 3)


def bardina(u, s, delta, g: Grid):
    """create_bardina's model (src/closures.jl:73-115): τ = G̃∗(u_i u_j) −
    (G̃∗u_i)(G̃∗u_j), test filter width sΔ, every filtered field 2/3-truncated,
    then make_tracefree! (physical)."""
    dt = s * delta
    ut = np.stack([twothirds(gaussianfilter(u[i], dt, g), g) for i in range(g.D)])
    sig = nonlinearity(u, g)
    sigt = nonlinearity(ut, g)
    sig = np.stack([twothirds(gaussianfilter(x, dt, g), g) for x in sig])
    tau = np.stack([to_phys(sig[i], g) - to_phys(sigt[i], g) for i in range(g.tensordim)])
    return make_tracefree(tau, g.D)


def smagorinsky_tensor(S, delta, D):
    """smagorinsky_tensor! (src/closures.jl:117-150); S packed."""
    if D == 2:
        S2 = S[0] ** 2 + S[1] ** 2 + 2 * S[2] ** 2
    else:
        S2 = S[0] ** 2 + S[1] ** 2 + S[2] ** 2 + 2 * S[3] ** 2 + 2 * S[4] ** 2 + 2 * S[5] ** 2
    nu = delta**2 * np.sqrt(2 * S2)
    return -2 * nu[None] * S


def ml_mm(M, L, D):
    """smagorinsky_ml_mm! (src/closures.jl:152-179): pointwise M:L_dev, M:M."""
    L = L.copy()
    if D == 2:
        tr = (L[0] + L[1]) / 2
        L[0] -= tr
        L[1] -= tr
        ml = M[0] * L[0] + M[1] * L[1] + 2 * M[2] * L[2]
        mm = M[0] * M[0] + M[1] * M[1] + 2 * M[2] * M[2]
        return ml, mm
    tr = (L[0] + L[1] + L[2]) / 3
    L[0] -= tr
    L[1] -= tr
    L[2] -= tr
    ml = M[0] * L[0] + M[1] * L[1] + M[2] * L[2] + 2 * M[3] * L[3] + 2 * M[4] * L[4] + 2 * M[5] * L[5]
    mm = M[0] * M[0] + M[1] * M[1] + M[2] * M[2] + 2 * M[3] * M[3] + 2 * M[4] * M[4] + 2 * M[5] * M[5]
    return ml, mm


def dynamic_coefficient(sum_ml, sum_mm):
    """c = max(-<ML>/<MM>, 0), 0 when <MM> == 0 (src/closures.jl:262)."""
    return 0.0 if sum_mm == 0 else max(-sum_ml / sum_mm, 0.0)


def dynamic_smagorinsky(u, delta, g: Grid, return_coeff=False, return_parts=False):
    """create_dynamic_smagorinsky's model (src/closures.jl:188-271) with the
    INTENDED test-filtered strain S̃ = to_phys(2/3 G̃ 2/3 Ŝ).  The reference
    computes S̃ from a buffer its own preceding C2R has destroyed
    (src/closures.jl:234-239); that value depends on FFT-library scratch and
    is not reproducible (parity unpinned against the reference's output).
    ``return_parts``: also return the intermediates L, m1 (dealiased), m2, M
    and c in a dict."""
    D = g.D
    dt = 2 * delta
    dd = np.sqrt(dt**2 + delta**2)
    ut = np.stack([twothirds(gaussianfilter(u[i], dt, g), g) for i in range(D)])
    sig = nonlinearity(u, g)
    sigt = nonlinearity(ut, g)
    sig = np.stack([twothirds(gaussianfilter(s, dt, g), g) for s in sig])
    L = np.stack([to_phys(sig[i], g) - to_phys(sigt[i], g) for i in range(g.tensordim)])
    Sh = strainrate(u, g)
    S, St = [], []
    for c in range(g.tensordim):
        sh = twothirds(Sh[c], g)
        S.append(to_phys(sh, g))
        St.append(to_phys(twothirds(gaussianfilter(sh, dt, g), g), g))
    S, St = np.stack(S), np.stack(St)
    m1 = smagorinsky_tensor(S, delta, D)
    m2 = smagorinsky_tensor(St, dd, D)
    m1 = np.stack([dealias_phys(m, g) for m in m1])
    M = np.stack([test_filter_phys(m, dt, g) for m in m1]) - m2
    ml, mm = ml_mm(M, L, D)
    c = dynamic_coefficient(float(np.sum(ml)), float(np.sum(mm)))
    tau = c * m1
    if return_parts:
        return tau, dict(L=L, m1=m1, m2=m2, M=M, c=c)
    return (tau, c) if return_coeff else tau


def gradient_phys(u, g: Grid):
    """∇ū in physical space: vectorgradient! + 2/3 + C2R per component
    (src/les.jl:12-16)."""
    G, _ = vectorgradient(u, g)
    return np.stack([to_phys(twothirds(G[i], g), g) for i in range(G.shape[0])])


def les_rhs(u, g: Grid, visc, model):
    """les! (src/les.jl:4-33).  ``model(u, GG) -> packed physical τ``."""
    sig = stress(u, visc, g)
    GG = gradient_phys(u, g)
    tau = model(u, GG)
    sig = sig + np.stack([to_spec(tau[i], g) for i in range(g.tensordim)])
    return tensordivergence(sig, g)


def contract_dissipation(tau, S, D):
    """contract_dissipation (src/solver.jl:746-759)."""
    if D == 2:
        return tau[0] * S[0] + tau[1] * S[1] + 2 * tau[2] * S[2]
    return (
        tau[0] * S[0] + tau[1] * S[1] + tau[2] * S[2]
        + 2 * (tau[3] * S[3] + tau[4] * S[4] + tau[5] * S[5])
    )


def apriori_metrics(tau_model, tau_ref):
    """relerr and crosscor of compute_sfs_stats (src/analysis.jl:116-121)."""
    a = np.asarray(tau_ref, float).ravel()
    b = np.asarray(tau_model, float).ravel()
    bb, aa = b - b.mean(), a - a.mean()
    relerr = np.linalg.norm(b - a) / np.linalg.norm(a)
    crosscor = np.dot(bb, aa) / np.sqrt(np.dot(bb, bb) * np.dot(aa, aa))
    return float(relerr), float(crosscor)


__all__ = [
    "clark", "smagorinsky", "dynamic_smagorinsky", "les_rhs", "gradient_phys",
    "contract_dissipation", "apriori_metrics", "make_tracefree", "ml_mm",
    "smagorinsky_tensor", "dynamic_coefficient", "grad_matrix", "pack_sym",
]
