"""Host-side mirror of src/filtering.jl and the SFS pair generator sfs!
(src/data.jl:614-642)."""
from ._lib import lib
from .grid import Grid, _p, ncomp_of
from .solver import _c, make_tracefree, nonlinearity, twothirds


def cutoff(ubar, u, g_les: Grid, g_dns: Grid):
    """``apply!(cutoff!, g_les, (ubar, u))`` (src/filtering.jl:28-34)."""
    _c(lib().sc_cutoff(g_les.ctx, g_dns.ctx, _p(ubar), _p(u), ncomp_of(u)), g_les)
    return ubar


def inverse_cutoff(u, ubar, g_les: Grid, g_dns: Grid):
    """inverse_cutoff! (src/filtering.jl:36-42)."""
    _c(lib().sc_inverse_cutoff(g_les.ctx, g_dns.ctx, _p(u), _p(ubar), ncomp_of(ubar)), g_les)
    return u


def gaussianfilter(u, delta, g: Grid):
    """gaussianfilter! (src/filtering.jl:44-67)."""
    _c(lib().sc_gaussianfilter(g.ctx, _p(u), ncomp_of(u), delta), g)
    return u


def sfs(u, g_dns: Grid, g_les: Grid, delta):
    """sfs (src/data.jl:614-623) / sfs! (:625-642): returns (ubar, tau) on the
    LES grid; ``delta=None`` skips the Gaussian filter."""
    sig = nonlinearity(g_dns.tensorfield(), u, g_dns)
    ubar = cutoff(g_les.vectorfield(), u, g_les, g_dns)
    if delta is not None:
        gaussianfilter(ubar, delta, g_les)
    twothirds(ubar, g_les)
    tau = cutoff(g_les.tensorfield(), sig, g_les, g_dns)  # sigma_bar_1
    del sig
    if delta is not None:
        gaussianfilter(tau, delta, g_les)
    s2 = nonlinearity(g_les.tensorfield(), ubar, g_les)  # sigma_bar_2
    axpy(tau, s2, -1.0, g_les)
    twothirds(tau, g_les)
    make_tracefree(tau, g_les)
    return ubar, tau


def axpy(y, x, a, g: Grid):
    """y .+= a .* x over all components (spectral)."""
    _c(lib().sc_axpy(g.ctx, _p(y), _p(x), a, ncomp_of(x)), g)
    return y
