"""Drivers for the benchmark configurations' reports (the reference runs its
studies through src/experiment.jl and the run-*.jl scripts; these are the
GPU-resident equivalents for the configurations this library targets).

* ``fp32_vs_fp64``: the 3D forced DNS in FP64 and FP32 from the same initial
  field and the same time steps; the relative spectrum difference, and the
  HBM rate of every FFT pass and of the whole RHS from the library's
  per-launch timers.
* ``pointwise_closures``: the pointwise learned closures trained on filtered
  DNS pairs; a-priori metrics, training and LES-step cost.
* ``les_classical``: forced DNS, Gaussian-filtered and cut to the LES grid;
  a-priori correlation of the dynamic Smagorinsky and Clark stresses with the
  exact sub-filter stress, then LES rollouts (FP32) with both closures and
  without one, against the filtered DNS at the same time.

Both return plain dicts (JSON-serialisable).
"""
import math
import time

import numpy as np
import torch

from . import solver as S
from .closures import create_clark, create_dynamic_smagorinsky, gradient_phys, les_rk4
from .filtering import sfs
from .grid import Grid
from .training import apriori_metrics, filter_reynolds

# model bytes per launch in units of one spectral scalar field C: the
# unpruned pass model of the fused RHS pipeline.  In the 3D RK4 step the last
# pass writes the divergence alone (z_fwd_div, 9 C) and the update adds the
# viscous and forcing terms (12, 18, 18, 15 C over the stages); pruned_c gives
# what the passes actually move, which the rocprofv3 FETCH/WRITE counters
# confirm
PASS_C = {"c2c_z_inv": 6, "c2c_y_inv": 6, "x_products": 9, "y_fwd_div": 12, "z_fwd_rhs": 12, "z_fwd_div": 9,
          "(k_rk4_update<T, 3>)": 13.5, "(k_rk4_update_lin<T, 3>)": 15.75}


def cast_spectral(u, g_from: Grid, g_to: Grid):
    """a spectral field copied to a grid of the same n in another precision"""
    out = g_to.spec(u.shape[0])
    out[..., :g_from.nh] = u[..., :g_from.nh].to(g_to.cdtype)
    return out


def spectrum_rel_diff(a, b):
    """relative L2 difference of two shell spectra (b the reference)"""
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def pruned_c(n):
    """bytes per launch (units of C) that the 3D passes actually move: the
    inverse passes read and write only the 2/3 band (kx <= n/3, then
    |ky| <= n/3, then |kz| <= n/3)"""
    kc, nh = n // 3, n // 2 + 1
    fx, fb = (kc + 1) / nh, (2 * kc + 1) / n
    return {"c2c_z_inv": 3 * fx * fb * fb + 3 * fx * fb, "c2c_y_inv": 3 * fx * fb + 3 * fx,
            "x_products": 3 * fx + 6, "y_fwd_div": 12, "z_fwd_rhs": 12, "z_fwd_div": 9,
            "(k_rk4_update<T, 3>)": 13.5, "(k_rk4_update_lin<T, 3>)": 15.75}


def _rates(g: Grid, rep, steps):
    C = (16 if g.prec == 64 else 8) * (g.n // 2 + 1) * g.n ** (g.D - 1)
    pc = pruned_c(g.n)
    out = {}
    for k, (cnt, ms) in rep.items():
        if k in PASS_C and cnt:
            t = ms / cnt / 1e3
            out[k] = {"ms_per_launch": ms / cnt, "GB_per_s": PASS_C[k] * C / t / 1e9,
                      "GB_per_s_moved": pc[k] * C / t / 1e9}
    rhs_keys = [k for k in ("c2c_z_inv", "c2c_y_inv", "x_products", "y_fwd_div", "z_fwd_rhs", "z_fwd_div") if k in rep]
    rhs_ms = sum(rep[k][1] for k in rhs_keys) / (4 * steps)
    rhs_c = sum(PASS_C[k] for k in rhs_keys)
    out["rhs"] = {"ms": rhs_ms, "GB_per_s": rhs_c * C / (rhs_ms / 1e3) / 1e9,
                  "GB_per_s_moved": sum(pc[k] for k in rhs_keys) * C / (rhs_ms / 1e3) / 1e9}
    return out


def fp32_vs_fp64(n=810, steps=100, visc=2.7e-4, cfl=0.3, seed=0, totalenergy=0.2, log=print):
    """3D forced DNS (band forcing 0.1 on |k| <= 2, E(k) ~ k^4 exp(-2(k/4)^2)),
    RK4 with the FP64 run's CFL step used by both runs.  ms_per_step leaves
    out the first step of each run, which carries the one-time choice of
    buffer placement."""
    g64 = Grid(3, n, dtype=torch.float64)
    g32 = Grid(3, n, dtype=torch.float32)
    u64 = S.randomfield(S.peak_profile, g64, seed=seed, totalenergy=totalenergy, kpeak=4)
    u32 = cast_spectral(u64, g64, g32)
    S.set_band_forcing(g64, 0.1, 2)
    S.set_band_forcing(g32, 0.1, 2)
    res = {"n": n, "steps": steps, "visc": visc, "cfl": cfl}
    _, e0 = S.spectrum(u64, g64)
    t = 0.0
    ms = {64: 0.0, 32: 0.0}
    S.profile(g64, True)
    S.profile(g32, True)
    for i in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dt = S.rk4_cfl(u64, g64, visc, cfl)
        torch.cuda.synchronize()
        if i > 0:
            ms[64] += (time.perf_counter() - t0) * 1e3
        t0 = time.perf_counter()
        S.rk4(u32, dt, g32, visc)
        torch.cuda.synchronize()
        if i > 0:
            ms[32] += (time.perf_counter() - t0) * 1e3
        t += dt
        if log and (i % 20 == 0 or i == steps - 1):
            log(f"step {i}: t={t:.4f} dt={dt:.3e}")
    rep64, rep32 = S.profile_report(g64), S.profile_report(g32)
    S.profile(g64, False)
    S.profile(g32, False)
    k, e64 = S.spectrum(u64, g64)
    _, e32 = S.spectrum(u32, g32)
    # pointwise relative difference over the shells holding energy (> 1e-10 of the peak)
    live = e64 > 1e-10 * e64.max()
    rel = np.where(live, np.abs(e32 - e64) / np.where(live, e64, 1.0), 0.0)
    res.update({
        "t_end": t,
        "energy": {"fp64": S.energy(u64, g64), "fp32": S.energy(u32, g32)},
        "spectrum_rel_l2": spectrum_rel_diff(e32, e64),
        "spectrum_rel_max": float(rel.max()),
        "spectrum_rel_max_k": float(k[int(rel.argmax())]),
        "spectrum_live_shells": int(live.sum()),
        "spectrum_change_fp64": spectrum_rel_diff(e64, e0),
        "ms_per_step": {"fp64": ms[64] / max(steps - 1, 1), "fp32": ms[32] / max(steps - 1, 1)},
        "rates_fp64": _rates(g64, rep64, steps),
        "rates_fp32": _rates(g32, rep32, steps),
    })
    return res


def les_classical(n_dns=512, n_les=128, visc=2e-3, factors=(2, 4), les_steps=500, spinup=1.0, cfl=0.3, seed=1,
                  totalenergy=0.2, log=print):
    """DNS (FP64) spin-up to t = spinup; a-priori stress metrics at that time;
    LES (FP32, RK4, fixed dt) for les_steps steps with dynamic Smagorinsky
    (test filter 2 Delta), Clark and no closure; the DNS advanced to the same
    time and filtered; LES spectrum errors against it."""
    g_dns = Grid(3, n_dns, dtype=torch.float64)
    g_les = Grid(3, n_les, dtype=torch.float64)
    g32 = Grid(3, n_les, dtype=torch.float32)
    S.set_band_forcing(g_dns, 0.1, 2)
    S.set_band_forcing(g32, 0.1, 2)
    u = S.randomfield(S.peak_profile, g_dns, seed=seed, totalenergy=totalenergy, kpeak=4)
    t = 0.0
    while t < spinup * (1 - 1e-12):
        t += S.rk4_cfl(u, g_dns, visc, cfl, spinup - t)
    res = {"n_dns": n_dns, "n_les": n_les, "visc": visc, "t0": t, "les_steps": les_steps, "filters": {}}
    # fixed LES step: half the DNS CFL step scaled to the coarser grid
    dt = 0.5 * cfl * S.propose_timestep(u, g_dns, visc) * (n_dns / n_les)
    res["dt_les"] = dt
    starts = {}
    for f in factors:
        delta = f * g_les.l / n_les
        ub, tau = sfs(u, g_dns, g_les, delta)
        tref = S.to_phys(g_les.spacetensorfield(), tau, g_les)
        tdyn = create_dynamic_smagorinsky(delta, g_les)(ub)
        tcl = S.make_tracefree(create_clark(delta, g_les)(ub, gradient_phys(ub, g_les)), g_les, spectral=False)
        res["filters"][f] = {
            "delta": delta, "redelta": filter_reynolds(ub, g_les, visc, delta),
            "apriori": {"dynsmag": apriori_metrics(tdyn, tref), "clark": apriori_metrics(tcl, tref)},
        }
        starts[f] = cast_spectral(ub, g_les, g32)
    # LES rollouts
    runs = {}
    for f in factors:
        delta = f * g_les.l / n_les
        for name in ("dynsmag", "clark", "none"):
            model = {"dynsmag": create_dynamic_smagorinsky(delta, g32), "clark": create_clark(delta, g32),
                     "none": None}[name]
            v = starts[f].clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ok = True
            for i in range(les_steps):
                les_rk4(v, dt, g32, model, visc)
                if i % 50 == 49:
                    e = S.energy(v, g32)
                    if not math.isfinite(e) or e > 1e5:
                        ok = False
                        break
            torch.cuda.synchronize()
            runs[(f, name)] = (v, ok, (time.perf_counter() - t0) / les_steps * 1e3)
            if log:
                log(f"LES Delta={f}h {name}: {runs[(f, name)][2]:.2f} ms/step, finite={ok}")
    # DNS to the LES end time
    T = t + les_steps * dt
    while t < T * (1 - 1e-12):
        t += S.rk4_cfl(u, g_dns, visc, cfl, T - t)
    res["t_end"] = t
    for f in factors:
        delta = f * g_les.l / n_les
        ub, _ = sfs(u, g_dns, g_les, delta)
        _, eref = S.spectrum(ub, g_les)
        out = {}
        for name in ("dynsmag", "clark", "none"):
            v, ok, msps = runs[(f, name)]
            _, e = S.spectrum(v, g32)
            out[name] = {"finite": ok, "ms_per_step": msps,
                         "spectrum_err": spectrum_rel_diff(e, eref) if ok else None}
        res["filters"][f]["aposteriori"] = out
    return res


def pointwise_closures(n_dns=256, n_les=64, visc=2e-3, factor=2, nsnap=6, dt_save=0.1, spinup=0.5, train_steps=300,
                       lr=3e-3, les_steps=20, cfl=0.3, seed=2, totalenergy=0.2, log=print):
    """The pointwise learned closures of the reference's study (mlp "conv",
    its group average "convsym", the TBNN and the equivariant net, at the
    p1200 / p3000 tiers) trained on the GPU DNS's own filtered pairs with the
    reference's semantics (create_dataloader / create_dataloader_tbnn pairs,
    relative loss, ClipNorm(1) + AdamW, warm-up + cosine; one snapshot per
    step); the a-priori relerr / crosscor of each on a held-out later snapshot
    against dynamic Smagorinsky and Clark; and the cost of one LES RK4 step
    (FP64, n_les^3) with each closure.  Returns a dict."""
    from . import pmlp as PM
    from .closures import create_smagorinsky
    from .training import create_data, reference_train

    g_dns = Grid(3, n_dns, dtype=torch.float64)
    g_les = Grid(3, n_les, dtype=torch.float64)
    S.set_band_forcing(g_dns, 0.1, 2)
    u = S.randomfield(S.peak_profile, g_dns, seed=seed, totalenergy=totalenergy, kpeak=4)
    t = 0.0
    while t < spinup * (1 - 1e-12):
        t += S.rk4_cfl(u, g_dns, visc, cfl, spinup - t)
    pairs = create_data(u, g_dns, g_les, visc, [factor], nsnap, dt_save, cfl=cfl)[factor]
    train, test = pairs[:-1], pairs[-1]
    delta = test.delta
    norm = PM.redelta_norm(pairs)
    nv = n_les**3
    tref = S.make_tracefree(S.to_phys(g_les.spacetensorfield(), test.tau, g_les), g_les, spectral=False)
    res = {"n_dns": n_dns, "n_les": n_les, "visc": visc, "delta": delta, "train_snapshots": len(train),
           "train_steps": train_steps, "redelta": [p.redelta for p in pairs], "nets": {}}

    def les_cost(model):
        v = test.ubar.clone()
        dt = 0.5 * cfl * S.propose_timestep(v, g_les, visc)
        les_rk4(v, dt, g_les, model, visc)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(les_steps):
            les_rk4(v, dt, g_les, model, visc)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / les_steps * 1e3

    specs = [("conv p3000", PM.PointwiseMLP, PM.ARCH_TIERS["conv"]["p3000"]),
             ("tbnn p3000", PM.TBNN, PM.ARCH_TIERS["tbnn"]["p3000"]),
             ("equi p1200", PM.EquivariantNet, PM.ARCH_TIERS["equi"]["p1200"])]
    for name, cls, widths in specs:
        net = cls(3, widths, use_redelta=True, seed=0, nvox_max=len(train) * nv)
        X, Y = PM.pointwise_pairs(net, train, g_les, norm)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        losses = reference_train(net, X, Y, train_steps, lr=lr, batchsize=nv)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / train_steps * 1e3
        del X, Y
        model = PM.PointwiseClosure(net, g_les, delta, visc, norm)
        r, c = apriori_metrics(S.make_tracefree(model(test.ubar), g_les, spectral=False), tref)
        out = {"widths": list(widths), "params": net.nparams, "ms_per_train_step": ms,
               "loss_first": losses[0], "loss_last": losses[-1], "apriori": {"relerr": r, "crosscor": c},
               "les_ms_per_step": les_cost(model)}
        if cls is PM.PointwiseMLP:
            sym = PM.symmetrize_pointwise(model)
            rs, cs = apriori_metrics(S.make_tracefree(sym(test.ubar), g_les, spectral=False), tref)
            res["nets"]["convsym p3000"] = {"apriori": {"relerr": rs, "crosscor": cs},
                                            "les_ms_per_step": les_cost(sym)}
        res["nets"][name] = out
        if log:
            log(f"{name}: {ms:.1f} ms/train step, loss {losses[0]:.3f} -> {losses[-1]:.3f}, "
                f"a-priori {r:.3f} / {c:.3f}, LES {out['les_ms_per_step']:.2f} ms/step")
    tdyn = create_dynamic_smagorinsky(delta, g_les)(test.ubar)
    tcl = S.make_tracefree(create_clark(delta, g_les)(test.ubar, gradient_phys(test.ubar, g_les)), g_les,
                           spectral=False)
    res["classical"] = {
        "dynsmag": {"apriori": dict(zip(("relerr", "crosscor"), apriori_metrics(tdyn, tref))),
                    "les_ms_per_step": les_cost(create_dynamic_smagorinsky(delta, g_les))},
        "clark": {"apriori": dict(zip(("relerr", "crosscor"), apriori_metrics(tcl, tref))),
                  "les_ms_per_step": les_cost(create_clark(delta, g_les))},
        "smagorinsky 0.17": {"les_ms_per_step": les_cost(create_smagorinsky(0.17, delta, g_les))},
        "none": {"les_ms_per_step": les_cost(None)},
    }
    return res


def pointwise_les_cost(n=64, steps=10, seed=1, visc=2e-3, dt=1e-3, log=print):
    """The cost of one FP64 LES RK4 step at n^3 with each pointwise closure at
    the reference's tiers (mlp and TBNN p3000, the mlp's group average,
    equivariant net p1200) on a random solenoidal field: the network weights
    are the initialisation (the cost does not depend on them).  Returns
    {closure: ms per RK4 step}."""
    from . import pmlp as PM

    g = Grid(3, n, dtype=torch.float64)
    u = S.randomfield(S.peak_profile, g, seed=seed, totalenergy=0.2, kpeak=4)
    delta = 2 * 2 * math.pi / n
    specs = [("mlp p3000", PM.PointwiseMLP, PM.ARCH_TIERS["conv"]["p3000"], False),
             ("mlp p3000 group average", PM.PointwiseMLP, PM.ARCH_TIERS["conv"]["p3000"], True),
             ("tbnn p3000", PM.TBNN, PM.ARCH_TIERS["tbnn"]["p3000"], False),
             ("equi p1200", PM.EquivariantNet, PM.ARCH_TIERS["equi"]["p1200"], False)]
    res = {"n": n, "steps": steps, "ms_per_step": {}}
    for name, cls, widths, sym in specs:
        net = cls(3, widths, use_redelta=True, seed=0, nvox_max=n**3)
        model = PM.PointwiseClosure(net, g, delta, visc, (0.5, 0.3), symmetrize=sym)
        v = u.clone()
        les_rk4(v, dt, g, model, visc)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            les_rk4(v, dt, g, model, visc)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / steps * 1e3
        res["ms_per_step"][name] = ms
        if log:
            log(f"{name}: {ms:.2f} ms per LES RK4 step at {n}^3")
    return res


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser()
    ap.add_argument("which", choices=("fp32_vs_fp64", "les_classical", "pointwise_closures", "pointwise_les_cost"))
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--steps", type=int, default=None)
    a = ap.parse_args()
    if a.which == "fp32_vs_fp64":
        kw = {k: v for k, v in (("n", a.n), ("steps", a.steps)) if v is not None}
        r = fp32_vs_fp64(**kw)
    elif a.which == "pointwise_les_cost":
        kw = {k: v for k, v in (("n", a.n), ("steps", a.steps)) if v is not None}
        r = pointwise_les_cost(**kw)
    elif a.which == "pointwise_closures":
        kw = {k: v for k, v in (("n_dns", a.n), ("train_steps", a.steps)) if v is not None}
        r = pointwise_closures(**kw)
    else:
        kw = {k: v for k, v in (("n_dns", a.n), ("les_steps", a.steps)) if v is not None}
        r = les_classical(**kw)
    print(json.dumps(r, default=str))
