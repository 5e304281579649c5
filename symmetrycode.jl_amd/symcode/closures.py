"""Host-side mirror of the classical closures (src/closures.jl) and the LES
right-hand side / rollout (src/les.jl).

A closure is a callable ``model(u, GG) -> tau`` like the reference's (packed
physical tensor, components xx, yy, zz, xy, yz, zx).  The built-in closures
also carry an id so that ``les_rhs`` can run the whole les! on the GPU in one
library call.
"""
import ctypes
import math

from ._lib import lib
from .grid import Grid, _p
from .solver import _c, energy, energy_shells, maintain_shell_energy, twothirds, wray3_cfl


class Closure:
    kind = 0
    p1 = 0.0
    p2 = 0.0

    def __init__(self, g: Grid):
        self.g = g


class Clark(Closure):
    """create_clark (src/closures.jl:64): τ = Δ²/12 · G Gᵀ."""
    kind = 1

    def __init__(self, delta, g):
        super().__init__(g)
        self.p1 = float(delta)

    def __call__(self, u, GG):
        tau = self.g.spacetensorfield()
        _c(lib().sc_closure_clark(self.g.ctx, _p(tau), _p(GG), self.p1), self.g)
        return tau


class Smagorinsky(Closure):
    """create_smagorinsky (src/closures.jl:65-66)."""
    kind = 2

    def __init__(self, cs, delta, g):
        super().__init__(g)
        self.p1, self.p2 = float(cs), float(delta)

    def __call__(self, u, GG):
        tau = self.g.spacetensorfield()
        _c(lib().sc_closure_smagorinsky(self.g.ctx, _p(tau), _p(GG), self.p1, self.p2), self.g)
        return tau


class DynamicSmagorinsky(Closure):
    """create_dynamic_smagorinsky (src/closures.jl:188-271); ``G`` is
    accepted for the shared interface but the model works from ``u``."""
    kind = 3

    def __init__(self, delta, g):
        super().__init__(g)
        self.p1 = float(delta)
        self.last_coefficient = None

    def __call__(self, u, GG=None):
        tau = self.g.spacetensorfield()
        c = ctypes.c_double()
        _c(lib().sc_closure_dynsmag(self.g.ctx, _p(tau), _p(u), self.p1, ctypes.byref(c)), self.g)
        self.last_coefficient = c.value
        return tau


class Verstappen(Closure):
    """create_verstappen (src/closures.jl:53-62,68-71): the Q-R model,
    ν = C²Δ²|r|/q with q = tr(S²)/2, r = tr(S³)/3; 3D only."""
    kind = 5

    def __init__(self, c, delta, g):
        if g.D != 3:
            raise ValueError("the Q-R model is only defined in 3D")
        super().__init__(g)
        self.p1, self.p2 = float(c), float(delta)

    def __call__(self, u, GG):
        tau = self.g.spacetensorfield()
        _c(lib().sc_closure_verstappen(self.g.ctx, _p(tau), _p(GG), self.p1, self.p2), self.g)
        return tau


class Bardina(Closure):
    """create_bardina (src/closures.jl:73-115): τ = G̃∗(u_i u_j) −
    (G̃∗u_i)(G̃∗u_j) made trace-free, test filter width s·Δ; works from u."""
    kind = 6

    def __init__(self, s, delta, g):
        super().__init__(g)
        self.p1, self.p2 = float(s), float(delta)

    def __call__(self, u, GG=None):
        tau = self.g.spacetensorfield()
        _c(lib().sc_closure_bardina(self.g.ctx, _p(tau), _p(u), self.p1, self.p2), self.g)
        return tau


class LearnedClosure(Closure):
    """The convolutional closure as an LES model: τ = s_out · net(ū / s_in,
    extra) with ū = C2R(2/3 û) (the solver-facing wrapper, like fullchain,
    src/nets.jl:311-334, but with global scales: s_in a velocity scale, s_out
    = Δ²⟨|∇ū|²⟩, extra the standardised log Re_Δ channel).  ``net`` is a
    ``cnn.ConvClosureNet`` with the grid's n, cin 3 or 4 and cout 6; it
    stays attached to the grid until another model is attached."""
    kind = 4

    def __init__(self, net, g, s_in, s_out, extra=0.0):
        super().__init__(g)
        self.net, self.s_in, self.s_out, self.extra = net, float(s_in), float(s_out), float(extra)
        self.attach()

    def attach(self):
        _c(lib().sc_set_cnn_closure(self.g.ctx, self.net.net, self.s_in, self.s_out, self.extra), self.g)

    def __call__(self, u, GG=None):
        tau = self.g.spacetensorfield()
        _c(lib().sc_closure_cnn(self.g.ctx, _p(tau), _p(u)), self.g)
        return tau


def create_learned_closure(net, g, s_in, s_out, extra=0.0):
    return LearnedClosure(net, g, s_in, s_out, extra)


def create_clark(delta, g):
    return Clark(delta, g)


def create_smagorinsky(cs, delta, g):
    return Smagorinsky(cs, delta, g)


def create_dynamic_smagorinsky(delta, g):
    return DynamicSmagorinsky(delta, g)


def create_verstappen(c, delta, g):
    return Verstappen(c, delta, g)


def create_bardina(s, delta, g):
    return Bardina(s, delta, g)


def gradient_phys(u, g: Grid):
    """∇ū in physical space as les! forms it (src/les.jl:11-16)."""
    GG = g.spacetensorfield_nonsym()
    _c(lib().sc_gradient_phys(g.ctx, _p(GG), _p(u)), g)
    return GG


def les_rhs(du, u, g: Grid, model, visc):
    """les!(du, u, grid, cache; model, visc) (src/les.jl:4-33)."""
    kind, p1, p2 = _kind(model)
    _c(lib().sc_les_rhs(g.ctx, _p(du), _p(u), visc, kind, p1, p2), g)
    return du


def les_wray3(u, dt, g: Grid, model, visc):
    """wray3!(les!, u, Δt, grid, cache; model, visc)."""
    kind, p1, p2 = _kind(model)
    _c(lib().sc_les_wray3(g.ctx, _p(u), dt, visc, kind, p1, p2), g)
    return u


def _kind(model):
    if model is None:
        return 0, 0.0, 0.0
    if hasattr(model, "attach"):
        model.attach()
    return model.kind, model.p1, model.p2


def les_rk4(u, dt, g: Grid, model, visc):
    """One RK4 step of les! (RK4 has no reference counterpart)."""
    kind, p1, p2 = _kind(model)
    _c(lib().sc_les_rk4(g.ctx, _p(u), dt, visc, kind, p1, p2), g)
    return u


def les_rk4_cfl(u, g: Grid, model, visc, cfl, dtmax):
    """RK4 step of les! with Δt = min(cfl · propose_timestep(u), dtmax); returns Δt."""
    kind, p1, p2 = _kind(model)
    dt = ctypes.c_double()
    _c(lib().sc_les_rk4_cfl(g.ctx, _p(u), cfl, visc, dtmax, kind, p1, p2, ctypes.byref(dt)), g)
    return dt.value


def les_rhs_vjp(out, u, w, g: Grid, model, visc):
    """out = J(u)^T w for les! with no closure or a LearnedClosure (whose
    parameter gradients accumulate), RFFT-weighted inner product."""
    kind, _, _ = _kind(model)
    _c(lib().sc_les_rhs_vjp(g.ctx, _p(out), _p(u), _p(w), visc, kind), g)
    return out


def les_rk4_vjp(u0, ubar, dt, g: Grid, model, visc, out=None):
    """(d les_rk4 step / d u0)^T ubar, stages recomputed from u0."""
    kind, _, _ = _kind(model)
    out = g.vectorfield() if out is None else out
    _c(lib().sc_les_rk4_vjp(g.ctx, _p(out), _p(u0), _p(ubar), dt, visc, kind), g)
    return out


def les_wray3_vjp(u0, ubar, dt, g: Grid, model, visc, out=None):
    kind, _, _ = _kind(model)
    out = g.vectorfield() if out is None else out
    _c(lib().sc_les_wray3_vjp(g.ctx, _p(out), _p(u0), _p(ubar), dt, visc, kind), g)
    return out


def les_wray3_cfl(u, g: Grid, model, visc, cfl, dtmax):
    """wray3!(les!) with Δt = min(cfl · propose_timestep(u), dtmax); returns Δt."""
    kind, p1, p2 = _kind(model)
    dt = ctypes.c_double()
    _c(lib().sc_les_wray3_cfl(g.ctx, _p(u), cfl, visc, dtmax, kind, p1, p2, ctypes.byref(dt)), g)
    return dt.value


def solve_les(u, times, g: Grid, visc, model, cfl, forced=True, onsnapshot=None):
    """solve_les! (src/les.jl:188-247): CFL-limited Wray3 rollout with the
    shell-energy clamp and the blow-up guard.  Returns the number of save
    times reached."""
    shells = energy_shells(g, [1, 2], u) if forced else None
    t = times[0]
    for i, tstop in enumerate(times):
        if i > 0:
            while t < tstop:
                if model is None:
                    dt = wray3_cfl(u, g, visc, cfl, tstop - t)
                else:
                    dt = les_wray3_cfl(u, g, model, visc, cfl, tstop - t)
                t += dt
                if forced:
                    maintain_shell_energy(u, shells, g)
                e = energy(u, g)
                if dt < 1.0e-8 or e > 1.0e5 or not math.isfinite(e):
                    return i
        twothirds(u, g)
        if onsnapshot is not None:
            onsnapshot(i, u)
    return len(times)
