#!/usr/bin/env python3
"""Benchmark: 810^3 forced-DNS RK4 steps/s on one MI355X (FP64 by default).

Workload (the 3D forced DNS configuration): 810^3 periodic grid, l = 2π,
random-phase solenoidal initial field with E(k) ∝ k^4 exp(-2 (k/4)^2)
(total energy 0.2), linear band forcing on |k| <= 2 at injection 0.1,
ν = 2.7e-4 (kmax·η >= 1.5 at ε = 0.1), classic RK4 with projected stages and
a CFL-0.3 time step recomputed every step.  One "step" is one RK4 step.
Fields are resident in HBM before the timed region starts.

Prints one JSON line (see the driver contract in the repository docs).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "symmetrycode.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

PEAK_HBM = 8.0e12  # B/s, MI355X HBM3E
METRIC = "810³ DNS RK4 steps/s + %HBM roofline; 128³ a-posteriori LES grad steps/s"

# Algorithmic HBM bytes per launch, in units of C (one spectral scalar field,
# 16·(n/2+1)·n² B in FP64), for the unpruned pass model of one RHS.
ALG_C = {
    "c2c_z_inv": 6, "c2c_y_inv": 6, "x_products": 9, "y_fwd_div": 12, "z_fwd_rhs": 12, "z_fwd_div": 9,
    "c2c_y_fwd": 12, "c2c_z_fwd": 12, "(k_div_visc<T, 3>)": 12,
    # RK4 update with the viscous and forcing terms: 12, 18, 18, 15 C over the stages
    "(k_rk4_update_lin<T, 3>)": 15.75,
    # last forward pass with the RK4 update fused (the pass model's 12 C plus
    # the update's 9 C in stages 0-2 and 6 C in stage 3): 21, 21, 21, 18 C
    # (the kernel moves 15, 21, 21, 15 C: no accumulator read at stage 0, and
    # the accumulator holds the running new state, so stage 3 reads no u0)
    "z_fwd_rk4": 20.25,
}
RK4_STEP_C = 213  # fused-pipeline model of one RK4 step (45 C per RHS + updates)
# HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE at
# 810^3 FP64, calibrated against a 4 GB device copy (FETCH x 2, WRITE x 1);
# profiles/rk4_810_fp64_counters.md holds the table and the commands
MEASURED_GB_810_FP64 = {"x_products": 34.15, "y_fwd_div": 52.13, "z_fwd_rk4": 79.46, "z_fwd_div": 38.87,
                        "(k_rk4_update_lin<T, 3>)": 67.35}


def cpu_baseline(n_sample=256, threads=None, reps=3):
    """Oracle (numpy / scipy.fft) RK4 step at n_sample^3 FP64 on the host
    cores: one warm-up step, then the median of ``reps`` timed steps,
    extrapolated to 810^3 with an N log N model (810^3 FP64 needs ~90 GB of
    host buffers)."""
    import numpy as np

    from oracle import spectral as S

    threads = threads or min(os.cpu_count() or 1, 16)
    S.WORKERS = threads
    g = S.Grid(3, n_sample)
    u = S.randomfield(lambda k: S.peak_profile(k, 4), g, seed=0, totalenergy=0.2)
    f = lambda v: S.convectiondiffusion(v, g, 2.7e-4, forcing=(0.1, 2))
    dt = 0.3 * S.propose_timestep(u, g, 2.7e-4)
    u = S.rk4(f, u, dt, g)  # warm-up
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        u = S.rk4(f, u, dt, g)
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    N1, N2 = 810.0**3, float(n_sample) ** 3
    scale = (N1 * math.log(N1)) / (N2 * math.log(N2))
    return {
        "value": 1.0 / (t * scale),
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"RK4 steps of the CPU oracle at {n_sample}^3 FP64 (scipy.fft workers={threads}): one "
                  f"warm-up step, then the median of {reps} ({', '.join(f'{x:.2f}' for x in ts)} s), "
                  f"extrapolated to 810^3 by N log N (x{scale:.1f})",
    }


PEAK_BF16 = 2.5e15  # FLOP/s, MI355X dense BF16 MFMA


def spectrum_error(u, ref, g):
    """relative L2 distance of the shell spectra over the resolved shells"""
    import numpy as np

    import symcode as sc

    _, a = sc.spectrum(u, g)
    _, b = sc.spectrum(ref, g)
    m = min(len(a), len(b), g.n // 3)
    a, b = np.asarray(a[:m]), np.asarray(b[:m])
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def bench_closure(train_steps=3, grad_steps=2, n_dns=512, n_les=128, nsteps=20, nroll=100, pretrain=150, log=print):
    """The learned-closure workloads.

    Data: a 512^3 FP64 forced DNS (nu = 2e-3, band forcing 0.1 on |k| <= 2,
    E(k) ~ k^4 exp(-2(k/4)^2)) filtered with Gaussian Delta in {2, 4} h_LES and
    cut to 128^3 at 8 save times: 16 (u_bar, tau) pairs.
    Training: one optimiser step on the batch of 16 with the reference's
    semantics = forward (trace-free output), relative loss, backward (data
    and weight gradients of all four 3x3x3 layers), ClipNorm(1) + AdamW;
    timed, then ``pretrain`` more steps (warm-up + cosine) before the
    a-posteriori part.
    A-posteriori: 128^3 FP32 LES with the trained closure from the filtered
    DNS, the filtered DNS advanced alongside as the reference trajectory.
    One gradient step = 20 RK4 steps + trajectory MSE + hand-written adjoint
    (stages recomputed) + Adam.  The closures' accuracy: shell-spectrum error
    and e_post against the filtered DNS after 20 and ``nroll`` steps."""
    import torch

    import symcode as sc
    from symcode import training as T

    visc, cfl = 2e-3, 0.3
    t_setup = time.perf_counter()
    g_dns = sc.Grid(3, n_dns, dtype=torch.float64)
    g_les = sc.Grid(3, n_les, dtype=torch.float64)
    u = sc.randomfield(sc.peak_profile, g_dns, seed=1, totalenergy=0.2, kpeak=4)
    sc.set_band_forcing(g_dns, 0.1, 2)
    T.create_data(u, g_dns, g_les, visc, [2], 1, 0.2, cfl=cfl)  # spin-up
    data = T.create_data(u, g_dns, g_les, visc, [2, 4], 8, 0.02, cfl=cfl)
    pairs = data[2] + data[4]
    nrm = T.Normalization.of(pairs, g_les)
    net = sc.cnn.ConvClosureNet(n_les, batch_max=16, cin=4, width=64, cout=6, seed=0, tracefree=True)
    step = lambda: net.train_step(x, y, 1e-3, loss="relative", clip=1.0)
    x, y = T.pack_pairs(net, g_les, pairs, nrm)
    log(f"closure data ready ({time.perf_counter() - t_setup:.1f} s)")
    # ---- a-priori training step, batch 16
    step()
    torch.cuda.synchronize()
    net.profile(True)
    t0 = time.perf_counter()
    for _ in range(train_steps):
        loss = step()
    torch.cuda.synchronize()
    t_train = (time.perf_counter() - t0) / train_steps
    rep = net.profile_report()
    net.profile(False)
    conv = {k: v for k, v in rep.items() if k.startswith("conv_")}
    conv_ms = sum(v[1] for v in conv.values()) / train_steps
    conv_flop = sum(v[2] for v in conv.values()) / train_steps
    nvox = 16 * n_les**3
    useful = 3 * 2.0 * nvox * 27 * (4 * 64 + 64 * 64 * 2 + 64 * 6)  # fwd + data grad + weight grad
    hist = T.reference_train(net, x, y, pretrain, lr=1e-3)
    cnn_train = {
        "value": 1.0 / t_train, "unit": "training steps/s", "ms_per_step": t_train * 1e3,
        "workload": f"3x3x3 circular CNN 4->64->64->64->6 (GELU), batch 16 x {n_les}^3 filtered "
                    f"{n_dns}^3 DNS pairs, BF16 MFMA / FP32 accumulate, trace-free output, relative loss, "
                    "ClipNorm(1) + AdamW lr 1e-3",
        "loss": loss,
        "pretrain": {"steps": pretrain, "relative_loss_first": hist[0], "relative_loss_last": hist[-1]},
        "conv_ms_per_step": conv_ms,
        "conv_tflops_executed": conv_flop / (conv_ms / 1e3) / 1e12,
        "mfma_util_executed": conv_flop / (conv_ms / 1e3) / PEAK_BF16,
        "mfma_util_useful": useful / (conv_ms / 1e3) / PEAK_BF16,
        "mfma_peak_tflops": PEAK_BF16 / 1e12,
        "kernels_ms_per_step": {k: v[1] / train_steps for k, v in sorted(rep.items(), key=lambda kv: -kv[1][1])},
    }
    log(f"cnn train: {cnn_train['ms_per_step']:.1f} ms/step, pretrain loss {hist[0]:.3f} -> {hist[-1]:.3f}")
    # ---- a-posteriori, 128^3 FP32
    g32 = sc.Grid(3, n_les, dtype=torch.float32)
    sc.set_band_forcing(g32, 0.1, 2)
    delta = 2 * g_les.l / n_les
    # fixed LES step: half the CFL-0.3 step of the DNS scaled to the 4x coarser grid
    dt = 0.5 * cfl * sc.propose_timestep(u, g_dns, visc) * (n_dns / n_les)
    ub, _ = sc.sfs(u, g_dns, g_les, delta)
    u0 = T.regrid(ub, g_les, g32)
    re0 = T.filter_reynolds(ub, g_les, visc, delta)
    refs = []
    for _ in range(max(nsteps, nroll)):
        # the DNS advances with sub-steps of its own CFL limit to each LES time
        tt = 0.0
        while tt < dt * (1 - 1e-12):
            tt += sc.rk4_cfl(u, g_dns, visc, cfl, dt - tt)
        refs.append(T.regrid(sc.sfs(u, g_dns, g_les, delta)[0], g_les, g32))
    del g_dns, u
    model = nrm.closure(net, g32, delta, re0)
    window = refs[:nsteps]
    theta = net.get_params()  # the rollouts below use the pre-trained network
    T.aposteriori_step(net, u0, window, dt, g32, visc, model, lr=1e-4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(grad_steps):
        Lc = T.aposteriori_step(net, u0, window, dt, g32, visc, model, lr=1e-4)
    torch.cuda.synchronize()
    t_grad = (time.perf_counter() - t0) / grad_steps
    net.set_params(theta)
    rollouts = {}
    ref_norm = sum(sc.spectraldot(r, r, g32).real for r in window) / (2 * nsteps)
    assert ref_norm > 0, "the filtered-DNS reference trajectory is zero"
    closures = (("learned", model), ("dynsmag", sc.create_dynamic_smagorinsky(delta, g32)), ("none", None),
                ("clark", sc.create_clark(delta, g32)))
    for name, m in closures:
        T.rollout_loss(u0, window, dt, g32, visc, m)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        Lm = T.rollout_loss(u0, window, dt, g32, visc, m)
        torch.cuda.synchronize()
        rate = 1.0 / (time.perf_counter() - t0)
        v = u0.clone()
        acc = {}
        for i, r in enumerate(refs[:nroll]):
            sc.les_rk4(v, dt, g32, m, visc)
            if i + 1 in (nsteps, nroll):
                acc[i + 1] = {"spectrum_err": spectrum_error(v, r, g32), "e_post": T.e_post(v, r, g32)}
        rollouts[name] = {"rollouts_per_s": rate, "rel_traj_mse": Lm / ref_norm, "after_steps": acc}
    sc.set_band_forcing(g32, 0.0)
    le, dy = rollouts["learned"]["after_steps"], rollouts["dynsmag"]["after_steps"]
    les_grad = {
        "value": 1.0 / t_grad, "unit": "gradient steps/s", "ms_per_grad_step": t_grad * 1e3,
        "workload": f"{n_les}^3 FP32 LES with the trained learned closure, {nsteps} RK4 steps (dt={dt:.3e}) from "
                    f"filtered {n_dns}^3 DNS, trajectory MSE, hand-written adjoint with per-step "
                    "checkpoints and stage recompute, Adam",
        "loss": Lc,
        "rollouts": rollouts,
        "spectrum_err_learned_vs_dynsmag": {k: [le[k]["spectrum_err"], dy[k]["spectrum_err"]] for k in le},
        "learned_at_most_dynsmag_spectrum_err": all(le[k]["spectrum_err"] <= dy[k]["spectrum_err"] for k in le),
        "grad_steps_vs_dynsmag_rollouts": (1.0 / t_grad) / rollouts["dynsmag"]["rollouts_per_s"],
    }
    log(f"les grad: {les_grad['ms_per_grad_step']:.1f} ms/step")
    return cnn_train, les_grad


def bench_cnn_only(steps, warmup, n=128, batch=16):
    """training steps of the learned closure (4->64->64->64->6, batch 16 x
    128^3) on seeded synthetic input and target, with the reference's
    semantics; per-kernel HIP-event times"""
    import torch

    import symcode as sc

    net = sc.cnn.ConvClosureNet(n, batch_max=batch, cin=4, width=64, cout=6, seed=0, tracefree=True)
    gen = torch.Generator(device="cuda").manual_seed(0)
    x = net.new_input(batch)
    x[..., :4] = torch.randn((batch, net.nvox, 4), generator=gen, device="cuda").to(torch.bfloat16)
    y = torch.randn((batch, net.nvox, net.cout), generator=gen, device="cuda")
    for _ in range(max(warmup, 1)):
        net.train_step(x, y, 1e-3, loss="relative", clip=1.0)
    torch.cuda.synchronize()
    net.profile(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        net.train_step(x, y, 1e-3, loss="relative", clip=1.0)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / steps
    rep = net.profile_report()
    net.profile(False)
    return {"workload": f"learned closure training step, batch {batch} x {n}^3, synthetic", "steps": steps,
            "ms_per_step": t * 1e3,
            "kernels_ms_per_step": {k: v[1] / steps for k, v in sorted(rep.items(), key=lambda kv: -kv[1][1])}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=810)
    ap.add_argument("--prec", type=int, default=64, choices=(32, 64))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=256)
    ap.add_argument("--no-closure", action="store_true",
                    help="skip the learned-closure training and a-posteriori gradient benchmarks")
    ap.add_argument("--cnn-only", action="store_true",
                    help="time only --steps training steps of the batch-16 x 128^3 learned closure on seeded "
                         "synthetic input (the kernels the counter runs look at)")
    ap.add_argument("--copy-calib-gb", type=float, default=0.0,
                    help="also run one device-to-device copy of this many GB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU slab/pencil sharding is "
                                   "not built (one GPU per box)", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    import symcode as sc

    if args.cnn_only:
        print(json.dumps(bench_cnn_only(args.steps, args.warmup)), flush=True)
        return
    n, nu, cfl = args.n, 2.7e-4, 0.3
    dtype = torch.float64 if args.prec == 64 else torch.float32
    g = sc.Grid(3, n, dtype=dtype, device=0)
    u = sc.randomfield(sc.peak_profile, g, seed=0, totalenergy=0.2, kpeak=4)
    sc.set_band_forcing(g, 0.1, 2)
    L = sc.lib()

    # K steps = one call of the multi-step entry point (the state stays in
    # the z-pencil-tiled layout between steps; the conversions of u into and
    # out of it are inside the timed region)
    if args.warmup > 0:
        sc.rk4_steps(u, g, nu, args.warmup, cfl=cfl)
    if args.copy_calib_gb > 0:
        nb = int(args.copy_calib_gb * 1e9) // 16
        a = torch.empty(nb, dtype=torch.complex128, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    torch.cuda.synchronize()
    L.sc_profile(g.ctx, 1)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    t0 = time.perf_counter()
    dts = sc.rk4_steps(u, g, nu, args.steps, cfl=cfl)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1)
    import ctypes

    buf = ctypes.create_string_buffer(1 << 16)
    L.sc_profile_report(g.ctx, buf, len(buf))
    L.sc_profile(g.ctx, 0)
    kern = {}
    for line in buf.value.decode().splitlines():
        name, cnt, tot = line.rsplit(" ", 2)
        kern[name] = (int(cnt), float(tot))
    esize = 16 if args.prec == 64 else 8
    C = esize * (n // 2 + 1) * n * n
    # dominant kernel: largest total time among those with a byte model
    dom = max((k for k in kern if k in ALG_C), key=lambda k: kern[k][1])
    cnt, tot = kern[dom]
    avg_s = tot / cnt / 1e3
    achieved = ALG_C[dom] * C / avg_s
    ms_step = ms / args.steps
    e = sc.energy(u, g)
    out = {
        "metric": METRIC,
        "value": 1e3 / ms_step,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.prec == 64 else "f32",
        "data": "synthetic (seeded random-phase solenoidal field, E(k) ~ k^4 exp(-2(k/4)^2))",
        "config": {"workload": f"3D forced DNS {n}^3, RK4, CFL 0.3, band forcing |k|<=2 at 0.1, "
                               f"nu=2.7e-4", "grid": n, "scheme": "rk4", "precision": args.prec},
        "roofline": {
            "bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9,
            "unit": "GB/s", "frac": achieved / PEAK_HBM,
            "traffic": MEASURED_GB_810_FP64.get(dom) if (n == 810 and args.prec == 64) else None,
            "traffic_unit": "GB per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated)",
            "alg_bytes_per_launch": ALG_C[dom] * C, "avg_launch_ms": avg_s * 1e3,
        },
        "step_roofline": {
            "alg_bytes_per_step": RK4_STEP_C * C,
            "frac": RK4_STEP_C * C / (ms_step / 1e3) / PEAK_HBM,
        },
        "kernels_ms_per_step": {k: v[1] / args.steps for k, v in sorted(kern.items(), key=lambda kv: -kv[1][1])},
        "energy_after": e,
        "dt_per_step": [float(d) for d in dts],
        "wall_s": wall,
    }
    if not args.no_closure:
        del u, g
        torch.cuda.empty_cache()
        out["cnn_train"], out["les_grad"] = bench_closure(log=lambda m: print(m, file=sys.stderr, flush=True))
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
