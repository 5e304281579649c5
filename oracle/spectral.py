"""Oracle: core spectral solver (restates src/solver.jl of the reference).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

Every function cites the reference ``file:line`` it restates.  Arrays are
``[z, y, x]`` C-order (identical memory order to the reference's Julia
``(x, y, z)`` column-major arrays).
"""
from __future__ import annotations

import numpy as np
import scipy.fft as sfft

from . import rng as _rng

WORKERS = 1  # scipy.fft worker threads; the CPU-baseline timing raises this.


class Grid:
    """``Grid{D}`` (src/solver.jl:131-151): periodic cube of side ``l``, ``n`` points."""

    def __init__(self, D: int, n: int, l: float = 2 * np.pi, dtype=np.float64):
        if n % 2 != 0:  # src/solver.jl:148
            raise AssertionError("Only even number of grid points supported.")
        if D not in (2, 3):
            raise ValueError("D must be 2 or 3")
        self.D, self.n, self.l = D, n, float(l)
        self.rdtype = np.dtype(dtype)
        self.cdtype = np.dtype(np.complex64 if self.rdtype == np.float32 else np.complex128)
        self.nh = n // 2 + 1
        ky = fftfreq_int(n)
        kx = np.arange(self.nh)
        if D == 3:
            self.kint = (kx[None, None, :], ky[None, :, None], ky[:, None, None])
        else:
            self.kint = (kx[None, :], ky[:, None])
        c = np.pi / self.l * 2  # `pi / g.l * 2 * k`  (src/solver.jl:178-183)
        self.kfull = tuple(c * k.astype(np.float64) for k in self.kint)

    # src/solver.jl:155-202
    @property
    def spacing(self):
        return self.l / self.n

    @property
    def tensordim(self):
        return 3 if self.D == 2 else 6

    @property
    def spec_shape(self):  # ndrange (src/solver.jl:197-198), reversed to [z,y,x]
        return (self.n,) * (self.D - 1) + (self.nh,)

    @property
    def phys_shape(self):
        return (self.n,) * self.D

    @property
    def fft_fac(self):  # get_fft_fac (src/solver.jl:207)
        return float(self.n**self.D)

    def k2int(self):
        """squared_wavenumber_int (src/solver.jl:185-192)."""
        return sum(k.astype(np.int64) ** 2 for k in self.kint) + np.zeros(self.spec_shape, np.int64)

    def k2full(self):
        """squared_wavenumber_full (src/solver.jl:194): ``(2π/l)² · k2int``."""
        return (np.pi / self.l * 2) ** 2 * self.k2int()

    def rfft_weight(self):
        """Weight 2 for 1 <= kx <= n/2-1, else 1: the reference's ``getenergy``
        doubling ``sum + sum(selectdim(.,1,2:kmax))`` (src/solver.jl:702-709)."""
        w = np.ones(self.spec_shape)
        w[..., 1 : self.nh - 1] = 2.0
        return w

    def zeros_spec(self, ncomp=None):
        shp = self.spec_shape if ncomp is None else (ncomp,) + self.spec_shape
        return np.zeros(shp, self.cdtype)


def fftfreq_int(n):
    """fftfreq_int (src/solver.jl:167): Nyquist index n/2 maps to -n/2."""
    j = np.arange(n)
    return np.where(j < (n + 1) >> 1, j, j - n)


# ---------------------------------------------------------------- transforms
def to_spec(phys, g: Grid):
    """to_spec! (src/solver.jl:214-215): unnormalised R2C divided by n^D."""
    s = sfft.rfftn(phys, axes=tuple(range(-g.D, 0)), workers=WORKERS)
    return (s / g.fft_fac).astype(g.cdtype)


def to_phys(spec, g: Grid):
    """to_phys! (src/solver.jl:210-211): normalised C2R times n^D, i.e. the
    unnormalised backward transform."""
    p = sfft.irfftn(spec, s=g.phys_shape, axes=tuple(range(-g.D, 0)), workers=WORKERS)
    return (p * g.fft_fac).astype(g.rdtype)


def rfft_unnormalised(phys, g: Grid):
    """``mul!(spec, plan, phys)`` (no scaling), as used by dealias_phys!."""
    return sfft.rfftn(phys, axes=tuple(range(-g.D, 0)), workers=WORKERS).astype(g.cdtype)


def irfft_normalised(spec, g: Grid):
    """``ldiv!(phys, plan, spec)`` (normalised inverse)."""
    return sfft.irfftn(spec, s=g.phys_shape, axes=tuple(range(-g.D, 0)), workers=WORKERS).astype(g.rdtype)


def dealias_phys(phys, g: Grid):
    """dealias_phys! (src/solver.jl:218-223)."""
    return irfft_normalised(twothirds(rfft_unnormalised(phys, g), g), g)


def test_filter_phys(phys, delta, g: Grid):
    """test_filter_phys! (src/solver.jl:226-232): Gaussian test filter then 2/3."""
    from .filtering import gaussianfilter

    s = rfft_unnormalised(phys, g)
    s = twothirds(gaussianfilter(s, delta, g), g)
    return irfft_normalised(s, g)


# ---------------------------------------------------------- spectral kernels
def twothirds_mask(g: Grid):
    """twothirds! mask (src/solver.jl:371-389): keep iff k != 0 and every
    |k_i| <= n ÷ 3."""
    kcut = g.n // 3
    keep = np.ones(g.spec_shape, bool)
    nonzero = np.zeros(g.spec_shape, bool)
    for k in g.kint:
        keep &= np.abs(k) <= kcut
        nonzero |= k != 0
    return keep & nonzero


def twothirds(u, g: Grid):
    return np.where(twothirds_mask(g), u, 0).astype(u.dtype)


def project(u, g: Grid):
    """project! (src/solver.jl:350-368): u -= k (k·u)/|k|², k=0 left intact."""
    k = g.kfull
    kdotu = sum(k[i] * u[i] for i in range(g.D))
    k2 = sum(k[i] * k[i] for i in range(g.D))
    with np.errstate(invalid="ignore", divide="ignore"):
        p = kdotu / k2
    p = np.where(k2 == 0, 0, p)  # the (1,1,1) index is the only k2 == 0 mode
    return np.stack([u[i] - k[i] * p for i in range(g.D)]).astype(u.dtype)


def symmetrize(u, g: Grid):
    """symmetrize! (src/solver.jl:398-430): Hermitian fix of the kx=0 and
    kx=n/2 planes (applied to one scalar component)."""
    n = g.n
    kmax = n // 2
    u = u.copy()
    if g.D == 2:
        for ix in (0, kmax):
            for iy in range(n):
                iy_self = iy == 0 or iy == kmax
                if iy_self:
                    u[iy, ix] = u[iy, ix].real
                elif iy > kmax:
                    u[iy, ix] = np.conj(u[n - iy, ix])
        return u
    for ix in (0, kmax):
        plane = u[:, :, ix].copy()  # [z, y]
        out = plane.copy()
        for iy in range(n):
            iy_self = iy == 0 or iy == kmax
            piy = iy if iy_self else n - iy
            for iz in range(n):
                iz_self = iz == 0 or iz == kmax
                piz = iz if iz_self else n - iz
                if iy_self and iz_self:
                    out[iz, iy] = plane[iz, iy].real
                elif (iy > piy) or (iy == piy and iz > piz):
                    # KA kernels read the array as it is being written; the
                    # source entries (iy < piy or the lower half) are never
                    # written themselves, so reading `plane` is equivalent.
                    out[iz, iy] = np.conj(plane[piz, piy])
        u[:, :, ix] = out
    return u


def vectordivergence(u, g: Grid):
    """vectordivergence! (src/solver.jl:484-493)."""
    return sum(1j * g.kfull[i] * u[i] for i in range(g.D))


def viscosity(sigma, u, visc, g: Grid):
    """viscosity! (src/solver.jl:458-476): σ_ij -= ν i (k_j u_i + k_i u_j)."""
    k = g.kfull
    s = sigma.copy()
    if g.D == 2:
        kx, ky = k
        ux, uy = u
        s[0] -= 2 * visc * 1j * kx * ux
        s[1] -= 2 * visc * 1j * ky * uy
        s[2] -= visc * 1j * (ky * ux + kx * uy)
        return s
    kx, ky, kz = k
    ux, uy, uz = u
    s[0] -= 2 * visc * 1j * kx * ux
    s[1] -= 2 * visc * 1j * ky * uy
    s[2] -= 2 * visc * 1j * kz * uz
    s[3] -= visc * 1j * (ky * ux + kx * uy)
    s[4] -= visc * 1j * (kz * uy + ky * uz)
    s[5] -= visc * 1j * (kx * uz + kz * ux)
    return s


def tensordivergence(sigma, g: Grid):
    """tensordivergence! (src/solver.jl:509-521): du_i = -i k_j σ_ij."""
    k = g.kfull
    if g.D == 2:
        kx, ky = k
        xx, yy, xy = sigma
        return np.stack([-1j * kx * xx - 1j * ky * xy, -1j * kx * xy - 1j * ky * yy])
    kx, ky, kz = k
    xx, yy, zz, xy, yz, zx = sigma
    return np.stack(
        [
            -1j * kx * xx - 1j * ky * xy - 1j * kz * zx,
            -1j * kx * xy - 1j * ky * yy - 1j * kz * yz,
            -1j * kx * zx - 1j * ky * yz - 1j * kz * zz,
        ]
    )


def vectorgradient(u, g: Grid):
    """vectorgradient! (src/solver.jl:711-733).  Returns the non-symmetric
    tensor in the reference's NamedTuple order: 3D
    ``[xx, yx, zx, xy, yy, zy, xz, yz, zz]`` (tensorfield_nonsym,
    src/solver.jl:264-276) where ``G.ij = i k_j u_i``."""
    k = g.kfull
    D = g.D
    G = {}
    names = "xyz"[:D]
    for i in range(D):
        for j in range(D):
            G[names[i] + names[j]] = 1j * k[j] * u[i]
    order = ["xx", "yx", "xy", "yy"] if D == 2 else ["xx", "yx", "zx", "xy", "yy", "zy", "xz", "yz", "zz"]
    return np.stack([G[o] for o in order]), order


def strainrate(u, g: Grid):
    """strainrate! (src/solver.jl:776-794)."""
    k = g.kfull
    if g.D == 2:
        kx, ky = k
        ux, uy = u
        return np.stack([1j * kx * ux, 1j * ky * uy, 1j * (ky * ux + kx * uy) / 2])
    kx, ky, kz = k
    ux, uy, uz = u
    return np.stack(
        [
            1j * kx * ux,
            1j * ky * uy,
            1j * kz * uz,
            1j * (ky * ux + kx * uy) / 2,
            1j * (kz * uy + ky * uz) / 2,
            1j * (kx * uz + kz * ux) / 2,
        ]
    )


# ------------------------------------------------------------------ the RHS
PAIRS3 = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0)]
PAIRS2 = [(0, 0), (1, 1), (0, 1)]


def pairs(D):
    return PAIRS2 if D == 2 else PAIRS3


def nonlinearity(u, g: Grid):
    """nonlinearity! (src/solver.jl:434-456): v_i = C2R(2/3 û_i)·nᴰ;
    σ_ij = R2C(v_i v_j)/nᴰ (σ itself is not truncated)."""
    v = [to_phys(twothirds(u[i], g), g) for i in range(g.D)]
    return np.stack([to_spec(v[i] * v[j], g) for (i, j) in pairs(g.D)])


def stress(u, visc, g: Grid):
    """stress! (src/solver.jl:478-482)."""
    return viscosity(nonlinearity(u, g), u, visc, g)


def convectiondiffusion(u, g: Grid, visc, forcing=None):
    """convectiondiffusion! (src/solver.jl:949-953).  ``forcing`` optionally
    adds the linear band forcing (new; see ``band_forcing``)."""
    du = tensordivergence(stress(u, visc, g), g)
    if forcing is not None:
        du = du + band_forcing(u, g, *forcing)
    return du


# -------------------------------------------------------------- time steps
WRAY3_A = (8 / 15, 5 / 12, 3 / 4)
WRAY3_B = (1 / 4, 0.0)


def wray3(f, u, dt, g: Grid):
    """wray3! (src/solver.jl:968-1012): 2N-storage RK3 with a projection after
    every stage.  Returns the new ``u``."""
    T = g.rdtype.type
    a = tuple(T(x) for x in WRAY3_A)
    b = tuple(T(x) for x in WRAY3_B)
    ustart = u.copy()
    u = u.copy()
    for i in range(3):
        du = f(u)
        if i != 0:
            u = ustart.copy()
        u = u + (a[i] * dt) * du  # axpy!(a[i]*Δt, du, u)
        u = project(u, g)
        if i < 2:
            ustart = ustart + (b[i] * dt) * du
    return u


def rk4(f, u, dt, g: Grid):
    """Classic 4-stage RK4 (new; the reference has no RK4).  Mirrors wray3!'s
    convention of projecting every stage value:
    k1=f(u0); u1=P(u0+dt/2 k1); k2=f(u1); u2=P(u0+dt/2 k2); k3=f(u2);
    u3=P(u0+dt k3); k4=f(u3); u=P(u0 + dt/6 (k1+2k2+2k3+k4))."""
    u0 = u.copy()
    k = f(u0)
    acc = k.copy()
    ui = project(u0 + (dt / 2) * k, g)
    k = f(ui)
    acc = acc + 2 * k
    ui = project(u0 + (dt / 2) * k, g)
    k = f(ui)
    acc = acc + 2 * k
    ui = project(u0 + dt * k, g)
    k = f(ui)
    acc = acc + k
    return project(u0 + (dt / 6) * acc, g)


def propose_timestep(u, g: Grid, visc):
    """propose_timestep (src/solver.jl:931-947): min(h/max|u|, h²/(2Dν))."""
    v = [to_phys(twothirds(u[i], g), g) for i in range(g.D)]
    vmax = np.sqrt(sum(vi**2 for vi in v)).max()
    h = g.l / g.n
    return min(h / vmax, h**2 / g.D / 2 / visc)


# ------------------------------------------------------------- reductions
def getenergy(u, g: Grid):
    """getenergy (src/solver.jl:702): RFFT-weighted Σ|û|² of one scalar."""
    return float(np.sum(g.rfft_weight() * np.abs(u) ** 2))


def energy(u, g: Grid):
    """energy (src/solver.jl:643-646)."""
    return sum(getenergy(u[i], g) for i in range(u.shape[0])) / 2


def spectraldot(u, v, g: Grid):
    """spectraldot (src/solver.jl:708-709): RFFT-weighted Σ conj(u) v."""
    return complex(np.sum(g.rfft_weight() * np.conj(u) * v))


def get_dissipation(u, visc, g: Grid):
    """get_dissipation! (src/solver.jl:796-811): ν Σ_w k² |û|²."""
    u2 = sum(np.abs(u[i]) ** 2 for i in range(g.D))
    return float(np.sum(g.rfft_weight() * visc * g.k2full() * u2))


def get_palinstrophy(u, visc, g: Grid):
    """get_palinstrophy! (src/solver.jl:813-827), 2D only."""
    u2 = sum(np.abs(u[i]) ** 2 for i in range(g.D))
    return float(np.sum(g.rfft_weight() * visc * g.k2full() ** 2 * u2))


def shell_of_modes(g: Grid):
    """Integer shell index floor(sqrt(k2int)) of every RFFT mode; shell i is
    ``i² <= k2int < (i+1)²`` (getshells, src/solver.jl:1080-1131)."""
    k2 = g.k2int()
    s = np.floor(np.sqrt(k2.astype(np.float64))).astype(np.int64)
    s = np.where((s + 1) ** 2 <= k2, s + 1, s)
    s = np.where(s**2 > k2, s - 1, s)
    return s


def spectrum(u, g: Grid):
    """spectrum with spectral_stuff's default shells 1..(2·(n/2))÷3
    (src/solver.jl:664-699).  Returns (k, s)."""
    kcut = (2 * (g.n // 2)) // 3
    sh = shell_of_modes(g)
    e = g.rfft_weight() * sum(np.abs(u[i]) ** 2 for i in range(u.shape[0])) / 2
    s = np.array([e[sh == i].sum() for i in range(1, kcut + 1)])
    return 2 * np.pi / g.l * np.arange(1, kcut + 1), s


def energy_shells(u, g: Grid, shells=(1, 2)):
    """energy_shells (src/solver.jl:1135-1141): reference energy per shell."""
    sh = shell_of_modes(g)
    w = g.rfft_weight()
    e = w * sum(np.abs(u[i]) ** 2 for i in range(u.shape[0])) / 2
    return [(i, float(e[sh == i].sum())) for i in shells]


def maintain_shell_energy(u, g: Grid, eshells):
    """maintain_shell_energy! (src/solver.jl:1146-1152)."""
    sh = shell_of_modes(g)
    w = g.rfft_weight()
    u = u.copy()
    for i, eref in eshells:
        m = sh == i
        eshell = float((w * sum(np.abs(u[c]) ** 2 for c in range(u.shape[0])))[m].sum() / 2)
        u[:, m] *= np.sqrt(eref / eshell)
    return u


def band_forcing(u, g: Grid, injection, kband=2):
    """Linear band forcing (new; nearest reference: the dead forced_rhs!,
    src/data.jl:4-23): f = A·u on 0 < |k_int| <= kband with
    A = injection / Σ_w,band |û|², so that Σ_w Re(conj(û)·f) = injection."""
    k2 = g.k2int()
    band = (k2 > 0) & (k2 <= kband * kband)
    w = g.rfft_weight()
    s = float(np.sum((w * sum(np.abs(u[i]) ** 2 for i in range(u.shape[0])))[band]))
    A = injection / s if s > 0 else 0.0
    return np.where(band[None], A * u, 0).astype(u.dtype)


def turbulence_statistics(u, visc, g: Grid):
    """turbulence_statistics (src/solver.jl:836-858, 3D).  Like the reference
    it truncates ``u``; the truncated field is returned with the stats."""
    u = np.stack([twothirds(u[i], g) for i in range(g.D)])
    D = g.D
    e = sum(getenergy(u[i], g) for i in range(D)) / 2
    diss = get_dissipation(u, visc, g)
    uavg = np.sqrt(2 * e / D)
    if D == 2:
        enstrophy = diss / (2 * visc)
        omega_rms = np.sqrt(2 * enstrophy)
        ens_diss = get_palinstrophy(u, visc, g)
        l_kra = (visc**3 / ens_diss) ** (1 / 6)
        l_omega = uavg / omega_rms
        return u, dict(
            e=e, uavg=uavg, diss=diss, enstrophy=enstrophy, palinstrophy=ens_diss / (2 * visc),
            enstrophy_diss=ens_diss, omega_rms=omega_rms, l_kra=l_kra, l_omega=l_omega,
            t_eddy=1 / omega_rms, t_ens=ens_diss ** (-1 / 3), Re=uavg * l_omega / visc,
            kmax_etaK=(np.pi * g.n / g.l) * l_kra,
        )
    l_kol = (visc**3 / diss) ** 0.25
    l_tay = np.sqrt(15 * visc / diss) * uavg
    l_int = uavg**3 / diss
    return u, dict(
        e=e, uavg=uavg, diss=diss, l_int=l_int, l_tay=l_tay, l_kol=l_kol,
        t_int=l_int / uavg, t_tay=l_tay / uavg, t_kol=np.sqrt(visc / diss),
        Re_int=l_int * uavg / visc, Re_tay=l_tay * uavg / visc,
        kmax_eta=(np.pi * g.n / g.l) * l_kol,
    )


# ------------------------------------------------------- initial conditions
def peak_profile(k, kpeak):
    """peak_profile (src/solver.jl:523)."""
    return k**4 * np.exp(-2 * (k / kpeak) ** 2)


def linear_profile(k, D):
    """linear_profile_2D / _3D (src/solver.jl:524-525)."""
    k = np.asarray(k, float)
    with np.errstate(divide="ignore"):
        return np.where(k > 0, k ** (-3.0 if D == 2 else -5.0 / 3.0), 0.0)


def random_spectral(g: Grid, seed: int, ncomp: int):
    """Seeded complex normals (re, im ~ N(0, 1/2)) on every RFFT mode; the
    counter-based stream replaces Julia's ``randn!`` (parity unpinned versus
    Xoshiro: the Julia stream is not reproducible outside Julia)."""
    return _rng.complex_normals(seed, ncomp, g.spec_shape)


def randomfield(profile, g: Grid, seed: int, totalenergy=1.0):
    """randomfield (src/solver.jl:576-641) with the counter-based normals."""
    u = random_spectral(g, seed, g.D).astype(g.cdtype)
    u = np.stack([symmetrize(u[i], g) for i in range(g.D)])
    u = project(u, g)
    u = np.stack([twothirds(u[i], g) for i in range(g.D)])
    E = sum(np.abs(u[i]) ** 2 for i in range(g.D)) / 2
    kmax = g.n // 2
    kdiag = int(np.floor(np.sqrt(g.D) * kmax))
    totalprofile = sum(float(profile(k)) for k in range(kdiag + 1))
    k2 = g.k2int()
    w = g.rfft_weight()
    for k in range(kdiag + 1):
        mask = (k * k <= k2) & (k2 < (k + 1) ** 2)
        Eshell = float(np.sum((w * E)[mask]))
        E0 = totalenergy * float(profile(k)) / totalprofile
        factor = np.sqrt(E0 / Eshell) if Eshell > 0 else 0.0
        u = np.where(mask[None], factor * u, u)
    u = np.stack([twothirds(u[i], g) for i in range(g.D)])
    return u.astype(g.cdtype)


def taylorgreen(g: Grid, V0=1.0, doproject=True):
    """taylorgreen (src/solver.jl:537-570) on cell centres (i-1/2)h."""
    n, l = g.n, g.l
    h = l / n
    x = np.linspace(h / 2, l - h / 2, n)
    if g.D == 2:
        X = x[None, :]
        Y = x[:, None]
        ux = -V0 * np.sin(np.pi * 2 * X / l) * np.cos(np.pi * 2 * Y / l)
        uy = V0 * np.cos(np.pi * 2 * X / l) * np.sin(np.pi * 2 * Y / l)
        u = np.stack([to_spec(ux, g), to_spec(uy, g)])
    else:
        X = x[None, None, :]
        Y = x[None, :, None]
        Z = x[:, None, None]
        ux = V0 * np.sin(np.pi * 2 * X / l) * np.cos(np.pi * 2 * Y / l) * np.cos(np.pi * 2 * Z / l)
        uy = -V0 * np.cos(np.pi * 2 * X / l) * np.sin(np.pi * 2 * Y / l) * np.cos(np.pi * 2 * Z / l)
        u = np.stack([to_spec(ux, g), to_spec(uy, g), np.zeros(g.spec_shape, g.cdtype)])
    return project(u, g) if doproject else u
