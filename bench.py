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
    "c2c_z_inv": 6, "c2c_y_inv": 6, "x_products": 9, "y_fwd_div": 12, "z_fwd_rhs": 12,
    "c2c_y_fwd": 12, "c2c_z_fwd": 12, "(k_div_visc<T, 3>)": 12,
}
RK4_STEP_C = 213  # fused-pipeline model of one RK4 step (45 C per RHS + updates)


def cpu_baseline(n_sample=256, threads=None):
    """Oracle (numpy / scipy.fft) RK4 step at n_sample^3 FP64 on the host
    cores, extrapolated to 810^3 with an N log N model."""
    import numpy as np

    from oracle import spectral as S

    threads = threads or min(os.cpu_count() or 1, 16)
    S.WORKERS = threads
    g = S.Grid(3, n_sample)
    u = S.randomfield(lambda k: S.peak_profile(k, 4), g, seed=0, totalenergy=0.2)
    f = lambda v: S.convectiondiffusion(v, g, 2.7e-4, forcing=(0.1, 2))
    dt = 0.3 * S.propose_timestep(u, g, 2.7e-4)
    t0 = time.perf_counter()
    S.rk4(f, u, dt, g)
    t = time.perf_counter() - t0
    N1, N2 = 810.0**3, float(n_sample) ** 3
    scale = (N1 * math.log(N1)) / (N2 * math.log(N2))
    return {
        "value": 1.0 / (t * scale),
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"one RK4 step of the CPU oracle at {n_sample}^3 FP64 ({t:.2f} s, scipy.fft "
                  f"workers={threads}), extrapolated to 810^3 by N log N (x{scale:.1f})",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=810)
    ap.add_argument("--prec", type=int, default=64, choices=(32, 64))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=256)
    ap.add_argument("--copy-calib-gb", type=float, default=0.0,
                    help="also run one device-to-device copy of this many GB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU slab/pencil sharding is "
                                   "design-only (see DESIGN.md)", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    import symcode as sc

    n, nu, cfl = args.n, 2.7e-4, 0.3
    dtype = torch.float64 if args.prec == 64 else torch.float32
    g = sc.Grid(3, n, dtype=dtype, device=0)
    u = sc.randomfield(sc.peak_profile, g, seed=0, totalenergy=0.2, kpeak=4)
    sc.set_band_forcing(g, 0.1, 2)
    L = sc.lib()

    def step():
        return sc.rk4_cfl(u, g, nu, cfl)

    for _ in range(args.warmup):
        step()
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
    for _ in range(args.steps):
        step()
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
            "unit": "GB/s", "frac": achieved / PEAK_HBM, "traffic": None,
            "alg_bytes_per_launch": ALG_C[dom] * C, "avg_launch_ms": avg_s * 1e3,
        },
        "step_roofline": {
            "alg_bytes_per_step": RK4_STEP_C * C,
            "frac": RK4_STEP_C * C / (ms_step / 1e3) / PEAK_HBM,
        },
        "kernels_ms_per_step": {k: v[1] / args.steps for k, v in sorted(kern.items(), key=lambda kv: -kv[1][1])},
        "energy_after": e,
        "wall_s": wall,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
