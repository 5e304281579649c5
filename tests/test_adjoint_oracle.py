"""Pins for the oracle adjoint (oracle/adjoint.py): dot-product tests against
central finite differences of the oracle RHS / steps, and an independent
torch (CPU, float64) autograd restatement of the RHS."""
import numpy as np
import pytest
import torch

from oracle import adjoint as A
from oracle import spectral as S


def field(g, seed, trunc=True, scale=1.0):
    r = np.random.default_rng(seed)
    u = np.stack([S.to_spec(r.standard_normal(g.phys_shape), g) for _ in range(g.D)])
    u = S.project(u, g)
    if trunc:
        u = np.stack([S.twothirds(c, g) for c in u])
    return scale * u


def dot(a, b, g):
    return sum(S.spectraldot(a[i], b[i], g) for i in range(a.shape[0])).real


@pytest.mark.parametrize("D,n", [(3, 12), (3, 16), (2, 32)])
def test_rhs_vjp_dot_product(D, n):
    """<w, J du> = <J^T w, du>; J du by central differences, exact for the
    quadratic RHS up to rounding."""
    g = S.Grid(D, n)
    u = field(g, 1)
    du = field(g, 2, trunc=False)  # directions with all modes
    w = np.stack([S.to_spec(np.random.default_rng(3 + i).standard_normal(g.phys_shape), g) for i in range(D)])
    nu, eps = 0.02, 1e-3
    f = lambda v: S.convectiondiffusion(v, g, nu)
    Jdu = (f(u + eps * du) - f(u - eps * du)) / (2 * eps)
    lhs = dot(w, Jdu, g)
    rhs = dot(A.rhs_vjp(u, w, g, nu), du, g)
    assert lhs == pytest.approx(rhs, rel=1e-11)


def test_rhs_vjp_with_band_forcing():
    g = S.Grid(3, 12)
    u = field(g, 4)
    du = field(g, 5)
    w = field(g, 6, trunc=False)
    nu, eps = 0.02, 1e-5
    f = lambda v: S.convectiondiffusion(v, g, nu, forcing=(0.1, 2))
    Jdu = (f(u + eps * du) - f(u - eps * du)) / (2 * eps)
    assert dot(w, Jdu, g) == pytest.approx(dot(A.rhs_vjp(u, w, g, nu, (0.1, 2)), du, g), rel=1e-8)


def _torch_rhs(ur, ui, g, nu):
    """convectiondiffusion! restated in torch (float64, CPU)."""
    D, n = g.D, g.n
    u = torch.complex(ur, ui)
    mask = torch.tensor(S.twothirds_mask(g))
    k = [torch.tensor(np.broadcast_to(kk, g.spec_shape).copy()) for kk in g.kfull]
    dims = tuple(range(-D, 0))
    fac = float(n**D)
    v = [torch.fft.irfftn(torch.where(mask, u[i], 0), s=g.phys_shape, dim=dims) * fac for i in range(D)]
    pairs = S.pairs(D)
    sig = [torch.fft.rfftn(v[i] * v[j], dim=dims) / fac for (i, j) in pairs]
    # viscosity!
    if D == 3:
        sig[0] = sig[0] - 2 * nu * 1j * k[0] * u[0]
        sig[1] = sig[1] - 2 * nu * 1j * k[1] * u[1]
        sig[2] = sig[2] - 2 * nu * 1j * k[2] * u[2]
        sig[3] = sig[3] - nu * 1j * (k[1] * u[0] + k[0] * u[1])
        sig[4] = sig[4] - nu * 1j * (k[2] * u[1] + k[1] * u[2])
        sig[5] = sig[5] - nu * 1j * (k[0] * u[2] + k[2] * u[0])
        xx, yy, zz, xy, yz, zx = sig
        return torch.stack([
            -1j * k[0] * xx - 1j * k[1] * xy - 1j * k[2] * zx,
            -1j * k[0] * xy - 1j * k[1] * yy - 1j * k[2] * yz,
            -1j * k[0] * zx - 1j * k[1] * yz - 1j * k[2] * zz,
        ])
    sig[0] = sig[0] - 2 * nu * 1j * k[0] * u[0]
    sig[1] = sig[1] - 2 * nu * 1j * k[1] * u[1]
    sig[2] = sig[2] - nu * 1j * (k[1] * u[0] + k[0] * u[1])
    xx, yy, xy = sig
    return torch.stack([-1j * k[0] * xx - 1j * k[1] * xy, -1j * k[0] * xy - 1j * k[1] * yy])


@pytest.mark.parametrize("D,n", [(3, 12), (2, 16)])
def test_rhs_vjp_matches_torch_autograd(D, n):
    g = S.Grid(D, n)
    u = field(g, 7, trunc=False)
    w = field(g, 8, trunc=False)
    nu = 0.03
    ur = torch.tensor(u.real.copy(), requires_grad=True)
    ui = torch.tensor(u.imag.copy(), requires_grad=True)
    f = _torch_rhs(ur, ui, g, nu)
    assert np.allclose(f.detach().numpy(), S.convectiondiffusion(u, g, nu), atol=1e-13)
    W = torch.tensor(g.rfft_weight())
    wt = torch.tensor(w)
    L = torch.sum(W * (torch.conj(wt) * f).real)
    gr, gi = torch.autograd.grad(L, (ur, ui))
    ref = (gr.numpy() + 1j * gi.numpy()) / g.rfft_weight()
    got = A.rhs_vjp(u, w, g, nu)
    assert np.linalg.norm(got - ref) < 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("scheme", ["rk4", "wray3"])
def test_step_vjp_dot_product(scheme):
    g = S.Grid(3, 12)
    u = field(g, 9)
    du = field(g, 10)
    ub = field(g, 11, trunc=False)
    nu, dt, eps = 0.02, 0.05, 1e-4
    f = lambda v: S.convectiondiffusion(v, g, nu)
    step = (lambda v: S.rk4(f, v, dt, g)) if scheme == "rk4" else (lambda v: S.wray3(f, v, dt, g))
    vjp = A.rk4_vjp if scheme == "rk4" else A.wray3_vjp
    Jdu = (step(u + eps * du) - step(u - eps * du)) / (2 * eps)
    assert dot(ub, Jdu, g) == pytest.approx(dot(vjp(u, ub, dt, g, nu), du, g), rel=1e-8)
