"""LES statistics reduced on the GPU: solve_les's per-snapshot reduce!
(src/les.jl:94-139) and compute_sfs_stats (src/analysis.jl:41-137).

Every field stays in HBM; only scalars and shell arrays come back.  Closures
are ``model(u, GG) -> packed physical tau``, None for no model ("nomo") or
the string "ref" for the filtered-DNS stress itself.
"""
import ctypes
import numpy as np
import torch

from ._lib import lib
from .closures import gradient_phys
from .grid import Grid, _p
from . import solver as S


def _c(rc, g):
    return S._c(rc, g)


def sfs_dissipation(tau, GG, g: Grid, out=None):
    """pointwise ε = -τ:S with S = sym(∇ū) (contract_dissipation,
    src/solver.jl:746-759); positive = drain"""
    out = g.spacescalarfield() if out is None else out
    _c(lib().sc_sfs_dissipation(g.ctx, _p(out), _p(tau), _p(GG)), g)
    return out


def shell_transfer(u, tau, g: Grid, nshell=None):
    """Σ over shell i of w Re(conj û · clo), clo = tensordivergence(2/3
    to_spec(τ)), shells 1..nshell (default: spectrum's 1..(2·(n/2))÷3)"""
    nshell = (2 * (g.n // 2)) // 3 if nshell is None else nshell
    out = (ctypes.c_double * nshell)()
    _c(lib().sc_shell_transfer(g.ctx, _p(u), _p(tau), nshell, out), g)
    return np.array(out[:])


def sample_mean(x, g: Grid):
    """FP64 mean of the device samples x"""
    x = x.contiguous()
    out = ctypes.c_double()
    _c(lib().sc_sample_mean(g.ctx, _p(x), x.numel(), ctypes.byref(out)), g)
    return out.value


def moments(x, g: Grid):
    """moments (src/analysis.jl:9-18) of the device samples x (any shape,
    the grid's precision) and the backscatter fraction mean(x < 0)"""
    x = x.contiguous()
    out = (ctypes.c_double * 5)()
    _c(lib().sc_moments(g.ctx, _p(x), x.numel(), out), g)
    return dict(zip(("mean", "median", "std", "skewness", "backscatter"), out[:]))


def apriori_metrics(tau_model, tau_ref, g: Grid):
    """(relerr, crosscor) over the stacked trace-free tensors
    (src/analysis.jl:116-121)"""
    a, b = tau_ref.contiguous(), tau_model.contiguous()
    if a.shape != b.shape:
        raise ValueError("tensor shapes differ")
    out = (ctypes.c_double * 2)()
    _c(lib().sc_apriori_metrics(g.ctx, _p(b), _p(a), a.numel(), out), g)
    return out[0], out[1]


def _tau_ref(tau_spec, g: Grid):
    """reference stress: 2/3-truncated, to physical space, trace-free"""
    t = tau_spec.clone()
    S.twothirds(t, g)
    phys = S.to_phys(g.spacetensorfield(), t, g)
    return S.make_tracefree(phys, g, spectral=False)


def compute_sfs_stats(pairs, g: Grid, model):
    """A-priori statistics of one closure over filtered-DNS pairs (objects with
    .ubar and .tau spectral on the LES grid), as compute_sfs_stats
    (src/analysis.jl:41-137) without the KDE: the model's stress is dealiased
    and made trace-free, ε_sfs = -τ:S is pooled over all snapshots for its
    moments and backscatter fraction, relerr and crosscor are averaged over
    snapshots ((0, 1) for "ref", (1, 0) for no model)."""
    np_ = g.phys_elems
    samples = torch.empty((len(pairs), np_), dtype=g.dtype, device=g.device)
    spec = g.spec(g.tensordim)
    rel, cor = [], []
    for i, p in enumerate(pairs):
        u = p.ubar
        GG = gradient_phys(u, g)
        tref = _tau_ref(p.tau, g)
        if model is None:
            tmod = torch.zeros_like(tref)
        elif isinstance(model, str) and model == "ref":
            tmod = tref.clone()
        else:
            tmod = model(u, GG).clone()
            S.dealias_phys(tmod, spec, g)
            S.make_tracefree(tmod, g, spectral=False)
        sfs_dissipation(tmod, GG, g, out=samples[i].view(1, g.nz, g.n, g.n))
        if model is None:
            rel.append(1.0)
            cor.append(0.0)
        elif isinstance(model, str):
            rel.append(0.0)
            cor.append(1.0)
        else:
            r, c = apriori_metrics(tmod, tref, g)
            rel.append(r)
            cor.append(c)
    m = moments(samples, g)
    return {"diss": m, "apriori": {"relerr": float(np.mean(rel)), "crosscor": float(np.mean(cor))}}


class LESReducer:
    """solve_les's reduce! (src/les.jl:94-139) as an onsnapshot callback:
    per snapshot e_post, kinetic energy, viscous dissipation, spectrum and the
    SFS drain, plus the shell-resolved transfer accumulated over snapshots
    (``transfer()`` returns (k, eps_sfs(k) = -T_accum / snapshots))."""

    def __init__(self, refs, g: Grid, visc, model, taus=None):
        self.refs, self.g, self.visc, self.model, self.taus = refs, g, visc, model, taus
        self.e_post, self.ke, self.eps_visc, self.eps_sfs, self.spectra = [], [], [], [], []
        self.k, _ = S.spectrum(g.vectorfield(), g)
        self.T_accum = np.zeros(len(self.k))

    def __call__(self, i, u):
        from .training import e_post

        g = self.g
        self.e_post.append(e_post(u, self.refs[i], g))
        self.ke.append(S.energy(u, g))
        self.eps_visc.append(S.get_dissipation(u, self.visc, g))
        self.spectra.append(S.spectrum(u, g)[1])
        if self.model is None:
            self.eps_sfs.append(0.0)
            return
        GG = gradient_phys(u, g)
        if isinstance(self.model, str) and self.model == "ref":
            tau = _tau_ref(self.taus[i], g)
        else:
            tau = self.model(u, GG).clone()
            S.make_tracefree(tau, g, spectral=False)
        eps = sfs_dissipation(tau, GG, g)
        self.eps_sfs.append(sample_mean(eps, g))  # -mean(τ:S)
        self.T_accum += shell_transfer(u, tau, g, len(self.k))

    def transfer(self):
        n = max(len(self.e_post), 1)
        return self.k, -self.T_accum / n
