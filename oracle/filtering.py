"""Oracle: spectral filter, coarsening and exact SFS pairs (restates
src/filtering.jl and src/data.jl:614-642).  TEST INFRASTRUCTURE ONLY -- see
``oracle/__init__.py``."""
import numpy as np

from .spectral import Grid, nonlinearity, twothirds, vectorgradient


def cutoff_index(nbar, n, i, is1):
    """cutoff_index (src/filtering.jl:9-13) on a 0-based index ``i``."""
    imax = nbar // 2 + (1 if is1 else 0)  # 1-based imax; i0 >= imax <=> i1 > imax
    return np.where(i >= imax, n - nbar + i, i)


def cutoff(u, g_les: Grid, g_dns: Grid):
    """cutoff! (src/filtering.jl:28-34): gather the coarse RFFT array from the
    fine one (negative wavenumbers counted from the end)."""
    nbar, n = g_les.n, g_dns.n
    ix = cutoff_index(nbar, n, np.arange(g_les.nh), True)
    iy = cutoff_index(nbar, n, np.arange(nbar), False)
    if g_les.D == 2:
        return u[np.ix_(iy, ix)]
    return u[np.ix_(iy, iy, ix)]


def gaussianfilter(u, delta, g: Grid):
    """gaussianfilter! (src/filtering.jl:44-67): û *= exp(-Δ²k²/24) except
    the forced shells |k_int|² < 9, which stay unfiltered."""
    k2 = g.k2full()
    w = np.exp(-(delta**2) * k2 / 24)
    w = np.where(g.k2int() < 9, 1.0, w)
    return (u * w).astype(u.dtype)


def make_tracefree(tau, D):
    """make_tracefree! (src/solver.jl:305-317)."""
    tau = tau.copy()
    if D == 2:
        tr = (tau[0] + tau[1]) / 2
        tau[0] -= tr
        tau[1] -= tr
    else:
        tr = (tau[0] + tau[1] + tau[2]) / 3
        tau[0] -= tr
        tau[1] -= tr
        tau[2] -= tr
    return tau


def sfs(u, g_dns: Grid, g_les: Grid, delta):
    """sfs! (src/data.jl:625-642).  Returns (ubar, tau) on the LES grid."""
    sig = nonlinearity(u, g_dns)
    ubar = []
    for c in range(g_dns.D):
        b = cutoff(u[c], g_les, g_dns)
        if delta is not None:
            b = gaussianfilter(b, delta, g_les)
        ubar.append(twothirds(b, g_les))
    ubar = np.stack(ubar)
    s1 = []
    for c in range(g_dns.tensordim):
        b = cutoff(sig[c], g_les, g_dns)
        if delta is not None:
            b = gaussianfilter(b, delta, g_les)
        s1.append(b)
    s1 = np.stack(s1)
    s2 = nonlinearity(ubar, g_les)
    tau = np.stack([twothirds(s1[i] - s2[i], g_les) for i in range(g_dns.tensordim)])
    return ubar, make_tracefree(tau, g_les.D)


def getenergy(u, g: Grid):
    """getenergy (src/solver.jl:701): Σ|û|² over the RFFT array with the
    interior x-modes counted twice."""
    return float(np.sum(g.rfft_weight() * np.abs(u) ** 2))


def filter_reynolds(ubar, g: Grid, visc, delta):
    """filter_reynolds (src/solver.jl:768-774): Re_Δ = Δ² √(Σ_ij
    getenergy(2/3(i k_j ū_i))) / ν.  The gradient is truncated to the 2/3
    box (and loses k = 0) whatever ū holds."""
    A, _ = vectorgradient(ubar, g)
    a2 = sum(getenergy(twothirds(a, g), g) for a in A)
    return delta**2 * np.sqrt(a2) / visc
