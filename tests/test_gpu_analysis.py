"""GPU parity of the LES statistics (solve_les's reduce!, src/les.jl:94-139;
compute_sfs_stats, src/analysis.jl:41-137) against oracle/analysis.py:
FP64 within 1e-12 relative, the median exactly."""
import numpy as np
import pytest
import torch

from helpers import oracle_solenoidal, print_case, rand_phys, rel
from oracle import analysis as OA
from oracle import closures as OC
from oracle import filtering as OF
from oracle import spectral as S

pytestmark = pytest.mark.gpu


def grids(D, n):
    import symcode as sc

    return S.Grid(D, n), sc.Grid(D, n, dtype=torch.float64)


@pytest.mark.parametrize("D,n", [(3, 16), (2, 32)])
def test_dissipation_transfer_parity(D, n):
    import symcode as sc
    from symcode import analysis as A

    print_case("sfs dissipation transfer", D, n)
    go, gg = grids(D, n)
    u = oracle_solenoidal(go, 41, trunc=True)
    tau = np.stack([rand_phys(go.phys_shape, 50 + i) for i in range(go.tensordim)])
    ud = gg.upload_spectral(u)
    GG = sc.gradient_phys(ud, gg)
    oGG = OC.gradient_phys(u, go)
    taud = gg.upload_physical(tau)
    eps = A.sfs_dissipation(taud, GG, gg)
    assert rel(gg.download_physical(eps)[0], OA.sfs_dissipation(tau, oGG, D)) < 1e-12
    T = A.shell_transfer(ud, taud, gg)
    To = OA.shell_transfer(u, tau, go)
    assert np.abs(T - To).max() < 1e-12 * np.abs(To).max()
    assert A.sample_mean(eps, gg) == pytest.approx(float(OA.sfs_dissipation(tau, oGG, D).mean()), rel=1e-12)


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("n", [2, 3, 1000, 1001, (1 << 20) + 3])
def test_median_radix_selection(n, prec):
    """The median's device radix selection against numpy on samples with
    repeated values, both signs and signed zeros (exact equality)."""
    import symcode as sc
    from symcode import analysis as A

    print_case("median", n, prec)
    dt = np.float64 if prec == 64 else np.float32
    gg = sc.Grid(3, 16, dtype=torch.float64 if prec == 64 else torch.float32)
    r = np.random.default_rng(n + prec)
    x = np.round(r.standard_normal(n) * 50) / 8  # many repeats
    x[: n // 7] = -0.0
    x[n // 7 : n // 5] = 0.0
    x = x.astype(dt)
    r.shuffle(x)
    m = A.moments(torch.tensor(x, device="cuda"), gg)
    xs = np.sort(x.astype(np.float64))
    ref = xs[(n - 1) // 2] if n % 2 else (xs[n // 2 - 1] + xs[n // 2]) / 2
    assert m["median"] == ref


@pytest.mark.parametrize("n", [1001, 4096])
def test_moments_and_apriori_parity(n):
    import symcode as sc
    from symcode import analysis as A

    print_case("moments apriori", n)
    gg = sc.Grid(3, 16, dtype=torch.float64)
    x = np.random.default_rng(n).standard_normal(n) ** 3 + 0.2
    m = A.moments(torch.tensor(x, device="cuda"), gg)
    mo = OA.moments(x)
    assert m["median"] == mo["median"]
    for k in ("mean", "std", "skewness", "backscatter"):
        assert m[k] == pytest.approx(mo[k], rel=1e-12, abs=1e-15), k
    a = np.random.default_rng(1).standard_normal((6, n))
    b = 0.7 * a + 0.3 * np.random.default_rng(2).standard_normal((6, n))
    r, c = A.apriori_metrics(torch.tensor(b, device="cuda"), torch.tensor(a, device="cuda"), gg)
    ro, co = OC.apriori_metrics(b, a)
    assert r == pytest.approx(ro, rel=1e-12) and c == pytest.approx(co, rel=1e-12)


def test_compute_sfs_stats_and_les_reducer():
    """compute_sfs_stats with Smagorinsky over two filtered-DNS pairs, and the
    reduce! statistics of the final state of a short LES rollout"""
    import symcode as sc
    from symcode import analysis as A
    from symcode.training import Pair

    print_case("compute_sfs_stats", 3, 32, 16)
    god, gol = S.Grid(3, 32), S.Grid(3, 16)
    gd = sc.Grid(3, 32, dtype=torch.float64)
    gl = sc.Grid(3, 16, dtype=torch.float64)
    delta = 2 * np.pi / 16 * 2
    pairs, opairs = [], []
    for seed in (61, 62):
        u = oracle_solenoidal(god, seed, trunc=True)
        ub, tau = sc.sfs(gd.upload_spectral(u), gd, gl, delta)
        pairs.append(Pair(ub, tau, delta, 1.0))
        opairs.append((gl.download_spectral(ub), gl.download_spectral(tau)))
    model = sc.create_smagorinsky(0.1, delta, gl)
    st = A.compute_sfs_stats(pairs, gl, model)
    so = OA.compute_sfs_stats(opairs, gol, lambda u_, G_: OC.smagorinsky(G_, 0.1, delta, 3))
    assert st["diss"]["median"] == pytest.approx(so["diss"]["median"], rel=1e-12)
    for k in ("mean", "std", "skewness", "backscatter"):
        assert st["diss"][k] == pytest.approx(so["diss"][k], rel=1e-11, abs=1e-15), k
    for k in ("relerr", "crosscor"):
        assert st["apriori"][k] == pytest.approx(so["apriori"][k], rel=1e-11), k
    # Smagorinsky with C_S > 0 only drains: no backscatter
    assert st["diss"]["backscatter"] == 0.0
    ref = A.compute_sfs_stats(pairs, gl, "ref")
    assert ref["apriori"] == {"relerr": 0.0, "crosscor": 1.0}
    # reduce! on a rollout
    refs = [p.ubar for p in pairs]
    red = A.LESReducer(refs, gl, 2e-3, model)
    u0 = pairs[0].ubar.clone()
    nred = sc.solve_les(u0, [0.0, 0.01], gl, 2e-3, model, 0.3, forced=False, onsnapshot=red)
    assert nred == 2 and len(red.e_post) == 2 and red.e_post[0] == 0.0
    uo = gl.download_spectral(u0)
    GG = OC.gradient_phys(uo, gol)
    tau = OF.make_tracefree(OC.smagorinsky(GG, 0.1, delta, 3), 3)
    assert red.eps_sfs[-1] == pytest.approx(float(OA.sfs_dissipation(tau, GG, 3).mean()), rel=1e-11)
    k, eps_k = red.transfer()
    assert eps_k.shape == k.shape and np.all(np.isfinite(eps_k))
    assert red.eps_sfs[-1] > 0
