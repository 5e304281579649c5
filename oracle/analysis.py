"""Oracle: the LES statistics of solve_les's reduce! (src/les.jl:94-139) and
compute_sfs_stats (src/analysis.jl:41-137).  TEST INFRASTRUCTURE ONLY -- see
oracle/__init__.py.

Pinned (tests/test_oracle.py): for an eddy-viscosity stress
tau = -nu_t (G + G^T) the dissipation is 2 nu_t S:S and the shell transfer is
-nu_t sum_shell w k^2 |u|^2 (solenoidal u); the moments against closed forms
of small samples.
"""
import numpy as np

from .closures import apriori_metrics, contract_dissipation, gradient_phys
from .filtering import make_tracefree
from .spectral import Grid, dealias_phys, shell_of_modes, tensordivergence, to_phys, to_spec, twothirds


def sfs_dissipation(tau, GG, D):
    """ε = -τ:S, S = sym(∇ū) (strain_from_gradient, src/solver.jl:742-744)"""
    from .closures import grad_matrix

    G = grad_matrix(GG, D)
    S = (G + np.swapaxes(G, 0, 1)) / 2
    Sp = np.stack([S[0, 0], S[1, 1], S[0, 1]]) if D == 2 else np.stack(
        [S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[1, 2], S[2, 0]])
    return -contract_dissipation(tau, Sp, D)


def shell_transfer(u, tau, g: Grid):
    """Σ over shell i = 1..(2·(n/2))÷3 of w Re(conj û · clo),
    clo = tensordivergence(2/3 to_spec(τ)) (src/les.jl:125-139)"""
    th = np.stack([twothirds(to_spec(t, g), g) for t in tau])
    clo = tensordivergence(th, g)
    T = g.rfft_weight() * sum((np.conj(u[i]) * clo[i]).real for i in range(g.D))
    kcut = (2 * (g.n // 2)) // 3
    sh = shell_of_modes(g)
    return np.array([T[sh == i].sum() for i in range(1, kcut + 1)])


def moments(x):
    """moments (src/analysis.jl:9-18) and the backscatter fraction"""
    x = np.asarray(x, np.float64).ravel()
    m = x.mean()
    s = x.std(ddof=1)
    skew = 0.0 if s == 0 else float(np.mean(((x - m) / s) ** 3))
    return {"mean": m, "median": float(np.median(x)), "std": s, "skewness": skew,
            "backscatter": float(np.mean(x < 0))}


def tau_ref(tau_spec, g: Grid):
    return make_tracefree(np.stack([to_phys(twothirds(t, g), g) for t in tau_spec]), g.D)


def compute_sfs_stats(pairs, g: Grid, model):
    """pairs: [(ubar, tau) spectral]; model(u, GG) -> packed physical τ"""
    samples, rel, cor = [], [], []
    for u, tau in pairs:
        GG = gradient_phys(u, g)
        tr = tau_ref(tau, g)
        tm = make_tracefree(np.stack([dealias_phys(t, g) for t in model(u, GG)]), g.D)
        samples.append(sfs_dissipation(tm, GG, g.D).ravel())
        r, c = apriori_metrics(tm, tr)
        rel.append(r)
        cor.append(c)
    return {"diss": moments(np.concatenate(samples)),
            "apriori": {"relerr": float(np.mean(rel)), "crosscor": float(np.mean(cor))}}
