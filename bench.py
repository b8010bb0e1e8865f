#!/usr/bin/env python3
"""Throughput benchmark of the device-resident projection step on one MI355X.

Workload: 3D periodic Taylor-Green on [0, 2pi]^3 at 256^3 (Re = 1600: viscosity 1/1600,
CFL 0.5, dt = final_time/100 for final_time = 100 * dt_upper), initial field
u = sin x cos y cos z, v = -cos x sin y cos z, w = 0, p = (cos2x + cos2y)(cos2z + 2)/16;
TVD (van Leer) advection, exact ADI Helmholtz predictor, MGPCG pressure solve at relative
tolerance 1e-10. Every step runs in both precisions: FP64 (the reference's arithmetic) and
MIXED (FP32 multigrid V-cycle inside the FP64 outer CG).

Prints one JSON line. `value` is mixed-precision steps/s; the FP64 figures, the
mixed-vs-FP64 field difference, the max |div u| over all steps, the roofline of the dominant
kernel (live HIP-event timing inside the timed region) and the reference CPU baseline
(timed on this host) are reported beside it.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--precision", choices=["both", "fp64", "mixed"], default="both")
    ap.add_argument("--profile-kernel", default="k_smooth12")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    return ap.parse_args()


# Algorithmic HBM bytes per launch of the finest-level kernels: (distinct arrays read +
# arrays written) x cells x element size; halo re-reads and caches excluded.
def kernel_bytes(name, n_cells, mixed):
    f64, lo = 8, (4 if mixed else 8)
    table = {
        # b (rhs) read, x2 written
        "k_smooth12": 2 * lo,
        # x2, b read; coarse b (1/8) written
        "k_residual_restrict": 2 * lo + lo / 8,
        # x2, b, coarse x (1/8) read; x4 written
        "k_prolong_smooth": 3 * lo + lo / 8,
        # x4, b read; x5 written
        "k_smooth_last": 3 * lo,
        # d read, q written
        "k_apply_dot": 2 * f64,
        # p, d, r, q read; p, r written (+ FP32 copy of r in mixed)
        "k_update": 6 * f64 + (4 if mixed else 0),
        # z, d read; d written
        "k_direction": 2 * f64 + lo,
    }
    return table[name] * n_cells


def cpu_baseline(n, nu, cfl, final_time, steps):
    """Times the reference's own CPU path (oracle/_ref/ref_golden, built from the reference
    sources) on this host, one thread; falls back to the oracle restatement if the reference
    binary is absent."""
    from oracle import refrun

    if refrun.available():
        t0 = time.perf_counter()
        _, meta = refrun.run("tg_steps", (), (), timeout=1800, nx=n, ny=n, nz=n, viscosity=nu, cfl=cfl,
                             final_time=final_time, tol=1e-10, maxit=400, steps=steps)
        wall = time.perf_counter() - t0
        el = meta.get("elapsed_seconds", wall)
        return {"value": steps / el, "unit": "steps/s", "cores": 1, "kind": "reference",
                "sample": f"{steps} step(s) of the same {n}^3 Taylor-Green case (reference-convention "
                          f"initial field, FP64, tol 1e-10), reference CPU sources built -O3 without "
                          f"fast-math, 1 thread; {el:.1f} s for the steps",
                "pressure_iterations": meta.get("pressure_iterations")}
    from oracle import ofref

    c = ofref.tg_config(n, n, n, nu, cfl, final_time, 400, 1e-10)
    s = ofref.tg_init(c, convention=1)
    t0 = time.perf_counter()
    ofref.tg_steps(c, s, steps)
    el = time.perf_counter() - t0
    return {"value": steps / el, "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": f"{steps} step(s) of the {n}^3 case with the oracle restatement, 1 thread; {el:.1f} s",
            "pressure_iterations": s.metrics.pressure_iterations}


def run_precision(sa, ctx, n, nu, dt, precision, warmup, steps, prof_kernel):
    import torch

    st = sa.Stepper(ctx, sa.TAYLOR_GREEN, n, n, n, nu, dt, 400, 1e-10, precision)
    st.initialize(convention=1)
    div_inf = 0.0
    for _ in range(warmup):
        m = st.run(1)
        div_inf = max(div_inf, m.divergence_linf)
    torch.cuda.synchronize()
    ctx.sync()
    ctx.profile_kernel(prof_kernel)
    iters = []
    t0 = time.perf_counter()
    for _ in range(steps):
        m = st.run(1)
        iters.append(m.pressure_iterations)
        div_inf = max(div_inf, m.divergence_linf)
    ctx.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = ctx.profile_read()
    ctx.profile_kernel("")
    vel, p, _, _, _ = st.download()
    st.close()
    return {"seconds": t1 - t0, "iters": iters, "div_linf": div_inf, "div_l2": m.divergence_l2,
            "prof": prof, "fields": vel + [p], "max_cfl": m.max_cfl}


def rel_l2(a, b):
    import numpy as np

    x = a[1:-1, 1:-1, 1:-1]
    y = b[1:-1, 1:-1, 1:-1]
    return float(np.sqrt(np.mean((x - y) ** 2)) / max(np.sqrt(np.mean(y * y)), 1e-300))


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU: the path is measured on one MI355X "
                                   "(multi-GPU sharding is a design note, not a measured configuration)",
                          "n_gpus": a.gpus}))
        return 2
    import torch

    import solver_amd as sa

    n = a.n
    nu = 1.0 / 1600.0
    cfl = 0.5
    L = 2 * math.pi
    dt_upper = cfl / (3 * n / L)
    final_time = 100 * dt_upper * (1 - 1e-14)
    dt, nsteps = sa.taylor_green_dt(n, n, n, cfl, final_time)
    assert nsteps == 100, nsteps
    ctx = sa.Context(0)
    precs = {"both": [sa.FP64, sa.MIXED], "fp64": [sa.FP64], "mixed": [sa.MIXED]}[a.precision]
    res = {}
    for prec in precs:
        res[prec] = run_precision(sa, ctx, n, nu, dt, prec, a.warmup, a.steps, a.profile_kernel)
        r = res[prec]
        print(f"# {'fp64' if prec == sa.FP64 else 'mixed'}: {a.steps / r['seconds']:.2f} steps/s, "
              f"{1e3 * r['seconds'] / a.steps:.2f} ms/step, MGPCG iterations/step "
              f"{sum(r['iters']) / len(r['iters']):.2f}, max |div u| {r['div_linf']:.2e}", flush=True)
    main_prec = sa.MIXED if sa.MIXED in res else sa.FP64
    r = res[main_prec]
    value = a.steps / r["seconds"]
    n_cells = n ** 3
    prof = r["prof"]
    avg_ms = prof["total_ms"] / max(prof["launches"], 1)
    nbytes = kernel_bytes(a.profile_kernel, n_cells, main_prec == sa.MIXED)
    achieved = nbytes / (avg_ms * 1e-3) / 1e9 if prof["launches"] else None
    out = {
        "metric": "3D TG 256^3 steps/s mixed vs FP64 at ||div u||_inf <= 1e-10",
        "value": value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * r["seconds"] / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 outer CG + f32 multigrid" if main_prec == sa.MIXED else "f64",
        "data": "synthetic (analytic Taylor-Green initial field)",
        "config": {"workload": f"3D periodic Taylor-Green {n}^3, Re=1600, CFL 0.5, 100-step dt, MGPCG tol 1e-10",
                   "nx": n, "ny": n, "nz": n, "dt": dt, "viscosity": nu, "parallelism": "single GPU"},
        "mgpcg_iterations_per_step": sum(r["iters"]) / len(r["iters"]),
        "max_div_linf": max(x["div_linf"] for x in res.values()),
        "roofline": {"bound": "hbm", "kernel": a.profile_kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": None,
                     "algorithmic_bytes_per_launch": nbytes, "avg_launch_ms": avg_ms,
                     "launches": prof["launches"]},
    }
    if sa.FP64 in res:
        f = res[sa.FP64]
        out["fp64"] = {"steps_per_s": a.steps / f["seconds"], "ms_per_step": 1e3 * f["seconds"] / a.steps,
                       "mgpcg_iterations_per_step": sum(f["iters"]) / len(f["iters"]), "max_div_linf": f["div_linf"]}
        if sa.MIXED in res:
            out["mixed_speedup_vs_fp64"] = value / out["fp64"]["steps_per_s"]
            out["mixed_vs_fp64_rel_l2"] = {k: rel_l2(x, y) for k, x, y in
                                           zip("uvwp", res[sa.MIXED]["fields"], res[sa.FP64]["fields"])}
    if not a.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(n, nu, cfl, final_time, a.cpu_steps)
        except Exception as e:  # reported, never fatal to the GPU measurement
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
