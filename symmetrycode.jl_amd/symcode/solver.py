"""Host-side mirror of the reference's core solver functions
(src/solver.jl).  Every function forwards to libsymcode_hip.so; the names
and argument meaning follow the reference (``!`` dropped; functions whose
reference form mutates an argument mutate the same argument here).
"""
import ctypes
import math

import numpy as np

from ._lib import check, lib
from .grid import Grid, _p, ncomp_of

_d = ctypes.c_double


def _c(rc, g):
    check(rc, g.ctx)


# ------------------------------------------------------------- transforms
def to_phys(phys, spec, g: Grid):
    """to_phys! (src/solver.jl:210-211): unnormalised C2R (spec kept intact)."""
    _c(lib().sc_to_phys(g.ctx, _p(phys), _p(spec), ncomp_of(spec)), g)
    return phys


def to_spec(spec, phys, g: Grid):
    """to_spec! (src/solver.jl:214-215): R2C / n^D."""
    _c(lib().sc_to_spec(g.ctx, _p(spec), _p(phys), ncomp_of(phys)), g)
    return spec


def irfft(phys, spec, g: Grid):
    """``ldiv!(phys, plan, spec)``: normalised C2R."""
    _c(lib().sc_irfft(g.ctx, _p(phys), _p(spec), ncomp_of(spec)), g)
    return phys


def rfft(spec, phys, g: Grid):
    """``mul!(spec, plan, phys)``: unnormalised R2C."""
    _c(lib().sc_rfft(g.ctx, _p(spec), _p(phys), ncomp_of(phys)), g)
    return spec


def dealias_phys(phys, spec, g: Grid):
    """dealias_phys! (src/solver.jl:218-223); ``spec`` is scratch."""
    rfft(spec, phys, g)
    twothirds(spec, g)
    return irfft(phys, spec, g)


def test_filter_phys(phys, spec, delta, g: Grid):
    """test_filter_phys! (src/solver.jl:226-232)."""
    from .filtering import gaussianfilter

    rfft(spec, phys, g)
    gaussianfilter(spec, delta, g)
    twothirds(spec, g)
    return irfft(phys, spec, g)


# ---------------------------------------------------------------- kernels
def twothirds(u, g: Grid):
    """twothirds! (src/solver.jl:371-389)."""
    _c(lib().sc_twothirds(g.ctx, _p(u), ncomp_of(u)), g)
    return u


def project(u, g: Grid):
    """project! (src/solver.jl:350-368)."""
    _c(lib().sc_project(g.ctx, _p(u)), g)
    return u


def symmetrize(u, g: Grid):
    """symmetrize! (src/solver.jl:398-430)."""
    _c(lib().sc_symmetrize(g.ctx, _p(u), ncomp_of(u)), g)
    return u


def viscosity(sigma, u, visc, g: Grid):
    """viscosity! (src/solver.jl:458-476)."""
    _c(lib().sc_viscosity(g.ctx, _p(sigma), _p(u), visc), g)
    return sigma


def tensordivergence(du, sigma, g: Grid):
    """tensordivergence! (src/solver.jl:509-521)."""
    _c(lib().sc_tensordivergence(g.ctx, _p(du), _p(sigma)), g)
    return du


def vectordivergence(div, u, g: Grid):
    """vectordivergence! (src/solver.jl:484-493)."""
    _c(lib().sc_vectordivergence(g.ctx, _p(div), _p(u)), g)
    return div


def vectorgradient(G, u, g: Grid):
    """vectorgradient! (src/solver.jl:711-733)."""
    _c(lib().sc_vectorgradient(g.ctx, _p(G), _p(u)), g)
    return G


def strainrate(S, u, g: Grid):
    """strainrate! (src/solver.jl:776-794)."""
    _c(lib().sc_strainrate(g.ctx, _p(S), _p(u)), g)
    return S


def make_tracefree(tau, g: Grid, spectral=True):
    """make_tracefree! (src/solver.jl:305-317)."""
    f = lib().sc_make_tracefree_spectral if spectral else lib().sc_make_tracefree_physical
    _c(f(g.ctx, _p(tau)), g)
    return tau


# ------------------------------------------------------------------ RHS
def nonlinearity(sigma, u, g: Grid):
    """nonlinearity! (src/solver.jl:434-456)."""
    _c(lib().sc_nonlinearity(g.ctx, _p(sigma), _p(u)), g)
    return sigma


def stress(sigma, u, visc, g: Grid):
    """stress! (src/solver.jl:478-482)."""
    _c(lib().sc_stress(g.ctx, _p(sigma), _p(u), visc), g)
    return sigma


def convectiondiffusion(du, u, g: Grid, visc):
    """convectiondiffusion!(du, u, grid, cache; visc) (src/solver.jl:949-953)."""
    _c(lib().sc_convectiondiffusion(g.ctx, _p(du), _p(u), visc), g)
    return du


def wray3(u, dt, g: Grid, visc):
    """wray3!(convectiondiffusion!, u, Δt, grid, cache; visc) (src/solver.jl:968-1012)."""
    _c(lib().sc_wray3(g.ctx, _p(u), dt, visc), g)
    return u


def rk4(u, dt, g: Grid, visc):
    """Classic RK4 with projected stages (no reference counterpart)."""
    _c(lib().sc_rk4(g.ctx, _p(u), dt, visc), g)
    return u


def rk4_cfl(u, g: Grid, visc, cfl, dtmax=float("inf")):
    """One RK4 step with Δt = min(cfl · propose_timestep(u), dtmax); the
    velocity maximum comes from the first stage's own C2R.  Returns Δt."""
    dt = _d()
    _c(lib().sc_rk4_cfl(g.ctx, _p(u), cfl, visc, min(dtmax, 1e300), ctypes.byref(dt)), g)
    return dt.value


def rk4_steps(u, g: Grid, visc, nsteps, cfl=0.0, dtmax=float("inf")):
    """nsteps RK4 steps in one call (the march loop of create_dns,
    src/data.jl:57-66, without its per-step host work): cfl > 0 takes
    min(cfl * propose_timestep, dtmax) every step, else dt = dtmax.  Returns
    the step sizes."""

    if cfl <= 0 and not math.isfinite(dtmax):
        raise ValueError("a fixed step needs a finite dtmax")
    dts = (ctypes.c_double * max(int(nsteps), 1))()
    _c(lib().sc_rk4_steps(g.ctx, _p(u), int(nsteps), float(cfl), visc, float(min(dtmax, 1e300)), dts), g)
    return np.array(dts[:nsteps])


def wray3_cfl(u, g: Grid, visc, cfl, dtmax=float("inf")):
    """wray3! with Δt = min(cfl · propose_timestep(u), dtmax) (the callers'
    step-size rule, src/data.jl:61-62).  Returns Δt."""
    dt = _d()
    _c(lib().sc_wray3_cfl(g.ctx, _p(u), cfl, visc, min(dtmax, 1e300), ctypes.byref(dt)), g)
    return dt.value


def propose_timestep(u, g: Grid, visc):
    """propose_timestep (src/solver.jl:931-947)."""
    dt = _d()
    _c(lib().sc_propose_timestep(g.ctx, _p(u), visc, ctypes.byref(dt)), g)
    return dt.value


def set_band_forcing(g: Grid, injection, kband=2):
    """Enable linear band forcing at a fixed injection rate inside the RHS
    (injection <= 0 disables)."""
    _c(lib().sc_set_band_forcing(g.ctx, injection, kband), g)


# ------------------------------------------------------------ reductions
def energy(u, g: Grid):
    """energy (src/solver.jl:643-646)."""
    e = _d()
    _c(lib().sc_energy(g.ctx, _p(u), ncomp_of(u), ctypes.byref(e)), g)
    return e.value


def getenergy(u, g: Grid):
    """getenergy (src/solver.jl:702) of the component(s) in ``u``."""
    return 2 * energy(u, g)


def spectraldot(u, v, g: Grid):
    """spectraldot (src/solver.jl:708-709), summed over components."""
    re, im = _d(), _d()
    _c(lib().sc_spectraldot(g.ctx, _p(u), _p(v), ncomp_of(u), ctypes.byref(re), ctypes.byref(im)), g)
    return complex(re.value, im.value)


def norm2(u, g: Grid):
    """Σ |û|² over the stored half spectrum without RFFT weights: the squared
    LinearAlgebra.norm of the reference's rfft arrays."""
    e = _d()
    _c(lib().sc_norm2(g.ctx, _p(u), ncomp_of(u), ctypes.byref(e)), g)
    return e.value


def get_dissipation(u, visc, g: Grid):
    """get_dissipation! (src/solver.jl:796-811)."""
    e = _d()
    _c(lib().sc_dissipation(g.ctx, _p(u), visc, ctypes.byref(e)), g)
    return e.value


def filter_reynolds(ubar, g: Grid, visc, delta):
    """filter_reynolds (src/solver.jl:768-774): Re_Δ = Δ² √(Σ_ij
    getenergy(2/3 (i k_j ū_i))) / ν."""
    e = _d()
    _c(lib().sc_filter_reynolds(g.ctx, _p(ubar), visc, delta, ctypes.byref(e)), g)
    return e.value


def get_palinstrophy(u, visc, g: Grid):
    """get_palinstrophy! (src/solver.jl:813-827)."""
    e = _d()
    _c(lib().sc_palinstrophy(g.ctx, _p(u), visc, ctypes.byref(e)), g)
    return e.value


def spectrum(u, g: Grid):
    """spectrum with the default shells 1..(2·(n/2))÷3 (src/solver.jl:664-699)."""
    kcut = (2 * (g.n // 2)) // 3
    s = (ctypes.c_double * kcut)()
    _c(lib().sc_spectrum(g.ctx, _p(u), ncomp_of(u), kcut, s), g)
    return 2 * np.pi / g.l * np.arange(1, kcut + 1), np.array(s[:])


def shell_energies(u, g: Grid, nshell):
    out = (ctypes.c_double * nshell)()
    _c(lib().sc_shell_energies(g.ctx, _p(u), ncomp_of(u), nshell, out), g)
    return np.array(out[:])


def energy_shells(g: Grid, shells, u):
    """energy_shells (src/solver.jl:1135-1141) for shells [1, ..., m]."""
    shells = list(shells)
    if shells != list(range(1, len(shells) + 1)):
        raise ValueError("shells must be 1..m")
    e = shell_energies(u, g, len(shells) + 1)
    return [(i, float(e[i])) for i in shells]


def maintain_shell_energy(u, shells, g: Grid):
    """maintain_shell_energy! (src/solver.jl:1146-1152)."""
    eref = (ctypes.c_double * len(shells))(*[e for _, e in shells])
    _c(lib().sc_maintain_shell_energy(g.ctx, _p(u), len(shells), eref), g)
    return u


def turbulence_statistics(u, visc, g: Grid):
    """turbulence_statistics (src/solver.jl:836-892).  Truncates ``u`` in
    place, like the reference."""
    twothirds(u, g)
    D = g.D
    e = energy(u, g)
    diss = get_dissipation(u, visc, g)
    uavg = math.sqrt(2 * e / D)
    if D == 2:
        enstrophy = diss / (2 * visc)
        omega_rms = math.sqrt(2 * enstrophy)
        ens_diss = get_palinstrophy(u, visc, g)
        l_kra = (visc**3 / ens_diss) ** (1 / 6)
        l_omega = uavg / omega_rms
        return dict(
            e=e, uavg=uavg, diss=diss, enstrophy=enstrophy, palinstrophy=ens_diss / (2 * visc),
            enstrophy_diss=ens_diss, omega_rms=omega_rms, l_kra=l_kra, l_omega=l_omega,
            t_eddy=1 / omega_rms, t_ens=ens_diss ** (-1 / 3), Re=uavg * l_omega / visc,
            kmax_etaK=(math.pi * g.n / g.l) * l_kra,
        )
    l_kol = (visc**3 / diss) ** 0.25
    l_tay = math.sqrt(15 * visc / diss) * uavg
    l_int = uavg**3 / diss
    return dict(
        e=e, uavg=uavg, diss=diss, l_int=l_int, l_tay=l_tay, l_kol=l_kol,
        t_int=l_int / uavg, t_tay=l_tay / uavg, t_kol=math.sqrt(visc / diss),
        Re_int=l_int * uavg / visc, Re_tay=l_tay * uavg / visc,
        kmax_eta=(math.pi * g.n / g.l) * l_kol,
    )


# -------------------------------------------------------- initial fields
def peak_profile(k, kpeak):
    """peak_profile (src/solver.jl:523)."""
    return k**4 * math.exp(-2 * (k / kpeak) ** 2)


def linear_profile_2D(k):
    return k ** (-3.0) if k > 0 else 0.0


def linear_profile_3D(k):
    return k ** (-5.0 / 3.0) if k > 0 else 0.0


def randomfield(profile, g: Grid, seed=0, totalenergy=1.0, **kwargs):
    """randomfield (src/solver.jl:576-641) with the library's counter-based
    normals keyed by ``seed``."""
    kdiag = int(math.floor(math.sqrt(g.D) * (g.n // 2)))
    prof = [float(profile(k, **kwargs)) for k in range(kdiag + 1)]
    tot = sum(prof)
    E0 = (ctypes.c_double * (kdiag + 1))(*[totalenergy * p / tot for p in prof])
    u = g.vectorfield()
    _c(lib().sc_randomfield(g.ctx, _p(u), seed, kdiag + 1, E0), g)
    return u


def randn(u, g: Grid, seed):
    _c(lib().sc_randn(g.ctx, _p(u), ncomp_of(u), seed), g)
    return u


def taylorgreen(g: Grid, V0=1.0, doproject=True):
    """taylorgreen (src/solver.jl:537-570)."""
    u = g.vectorfield()
    _c(lib().sc_taylorgreen(g.ctx, _p(u), V0, 1 if doproject else 0), g)
    return u


# ------------------------------------------------------------------ adjoint
def rhs_vjp(out, u, w, g: Grid, visc):
    """out = J(u)^T w for convectiondiffusion! (plus band forcing when set),
    under the RFFT-weighted inner product of spectraldot."""
    _c(lib().sc_rhs_vjp(g.ctx, _p(out), _p(u), _p(w), visc), g)
    return out


def rk4_vjp(u0, ubar, dt, g: Grid, visc):
    """(d rk4 step / d u0)^T ubar for one step of fixed dt; returns a new field."""
    out = g.vectorfield()
    _c(lib().sc_rk4_vjp(g.ctx, _p(out), _p(u0), _p(ubar), dt, visc), g)
    return out


def wray3_vjp(u0, ubar, dt, g: Grid, visc):
    """(d wray3! step / d u0)^T ubar for one step of fixed dt."""
    out = g.vectorfield()
    _c(lib().sc_wray3_vjp(g.ctx, _p(out), _p(u0), _p(ubar), dt, visc), g)
    return out


def lincomb(dst, x, a, y, b, g: Grid):
    """dst = a x + b y over all components (spectral; dst may alias x or y)."""
    _c(lib().sc_lincomb(g.ctx, _p(dst), _p(x), a, _p(y), b, ncomp_of(x)), g)
    return dst


def profile(g: Grid, enable=True):
    """start (clearing the record) or stop per-launch HIP-event timing of the
    library's kernels on this grid's context"""
    _c(lib().sc_profile(g.ctx, 1 if enable else 0), g)


def profile_report(g: Grid):
    """{kernel: (launches, total ms)} recorded since profile(g, True)"""
    buf = ctypes.create_string_buffer(1 << 16)
    _c(lib().sc_profile_report(g.ctx, buf, len(buf)), g)
    out = {}
    for line in buf.value.decode().splitlines():
        name, cnt, tot = line.rsplit(" ", 2)
        out[name] = (int(cnt), float(tot))
    return out
