"""GPU parity of sc_rk4_steps, the multi-step RK4 entry point whose 3D path
keeps the Runge-Kutta state in z-pencil tiles and fuses the stage update,
viscous term, forcing and projection into the last forward pass: against the
oracle's RK4 with the CFL step (small grids, FP64 at 1e-12 per step, FP32 at
2e-5), against the per-step entry point at 810^3, and the layout round trip."""
import numpy as np
import pytest
import torch

from helpers import oracle_solenoidal, print_case, rel
from oracle import spectral as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("D,n,prec", [(3, 16, 64), (3, 30, 64), (3, 48, 64), (3, 30, 32), (2, 64, 64), (3, 512, 64)])
def test_rk4_steps_against_oracle(D, n, prec):
    """three CFL steps in one call, band forcing on; (3, 512, 64) checks the
    per-step fallback where the fused stage does not fit (first step only,
    against the per-step entry point)"""
    import symcode as sc

    print_case("rk4_steps", D, n, prec)
    go = S.Grid(D, n)
    gg = sc.Grid(D, n, dtype=torch.float64 if prec == 64 else torch.float32)
    nu, cfl = 0.01, 0.3
    sc.set_band_forcing(gg, 0.1, 2)
    if n == 512:
        u = sc.randomfield(sc.peak_profile, gg, seed=2, totalenergy=0.2, kpeak=4)
        v = u.clone()
        d1 = sc.rk4_steps(u, gg, nu, 1, cfl=cfl)
        d2 = sc.rk4_cfl(v, gg, nu, cfl)
        assert d1[0] == d2
        assert torch.equal(u, v)
        sc.set_band_forcing(gg, 0.0)
        return
    u = oracle_solenoidal(go, 51, trunc=True)
    ud = gg.upload_spectral(u)
    dts = sc.rk4_steps(ud, gg, nu, 3, cfl=cfl)
    f = lambda w: S.convectiondiffusion(w, go, nu, forcing=(0.1, 2))
    ref = u
    for k in range(3):
        dto = cfl * S.propose_timestep(ref, go, nu)
        assert dts[k] == pytest.approx(dto, rel=1e-12 if prec == 64 else 1e-5)
        ref = S.rk4(f, ref, dts[k], go)
    e = rel(gg.download_spectral(ud), ref)
    print(f"rk4_steps {D} {n} fp{prec}: rel {e:.2e}")
    assert e < (3e-12 if prec == 64 else 6e-5)
    # fixed dt (cfl <= 0)
    ud = gg.upload_spectral(u)
    sc.rk4_steps(ud, gg, nu, 2, dtmax=0.01)
    ref = S.rk4(f, S.rk4(f, u, 0.01, go), 0.01, go)
    assert rel(gg.download_spectral(ud), ref) < (2e-12 if prec == 64 else 4e-5)
    sc.set_band_forcing(gg, 0.0)


@pytest.mark.parametrize("prec", [64, 32])
def test_rk4_steps_810_against_per_step(prec):
    """810^3 (the fused stage at the benchmark size): two CFL steps in one
    call against two calls of the per-step entry point (the unfused update),
    band forcing on, and nsteps = 0 (conversion in and out) exact."""
    import symcode as sc

    print_case("rk4_steps 810", prec)
    g = sc.Grid(3, 810, dtype=torch.float64 if prec == 64 else torch.float32)
    u = sc.randomfield(sc.peak_profile, g, seed=0, totalenergy=0.2, kpeak=4)
    sc.set_band_forcing(g, 0.1, 2)
    nu, cfl = 2.7e-4, 0.3
    v = u.clone()
    sc.rk4_steps(v, g, nu, 0, cfl=cfl)
    assert torch.equal(u, v)
    d_fused = sc.rk4_steps(v, g, nu, 2, cfl=cfl)
    d_step = [sc.rk4_cfl(u, g, nu, cfl) for _ in range(2)]
    print(f"810 fp{prec}: dt fused {d_fused} per-step {d_step}")
    assert np.allclose(d_fused, d_step, rtol=1e-13 if prec == 64 else 1e-6, atol=0)
    du = u - v
    e = float(torch.linalg.vector_norm(du) / torch.linalg.vector_norm(u))
    print(f"810 fp{prec}: fused vs per-step rel {e:.2e}")
    assert e < (1e-13 if prec == 64 else 1e-5)
    sc.set_band_forcing(g, 0.0)
