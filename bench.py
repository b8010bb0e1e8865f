#!/usr/bin/env python3
"""Throughput benchmark of the device-resident projection step on one MI355X.

Workload: 3D periodic Taylor-Green on [0, 2pi]^3 at 256^3 (Re = 1600: viscosity 1/1600,
CFL 0.5, dt = final_time/100 for final_time = 100 * dt_upper), initial field
u = sin x cos y cos z, v = -cos x sin y cos z, w = 0, p = (cos2x + cos2y)(cos2z + 2)/16;
TVD (van Leer) advection, exact ADI Helmholtz predictor, MGPCG pressure solve at relative
tolerance 1e-10. Every step runs in both precisions: FP64 (the reference's arithmetic) and
MIXED (FP32 multigrid V-cycle inside the FP64 outer CG).

Prints one JSON line. `value` is mixed-precision steps/s; beside it: the FP64 figures, the
mixed-vs-FP64 field difference, the max |div u| over all steps, the whole-step HBM rate from
the algorithmic byte model, the roofline of the dominant kernel (live HIP-event timing of its
launches inside the timed region, on the stream they run on) with the HBM traffic measured by
rocprofv3 counters (profiles/), and the reference CPU path timed on this host.

--copy-calibration runs a plain 1 GiB device-to-device copy three times and exits: the known
byte count that scales the FETCH_SIZE / WRITE_SIZE counters.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "3D TG 256³ steps/s mixed vs FP64 at ‖div u‖∞≤1e-10; HBM GB/s; adjoint s/step"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
TRAFFIC_TABLE = os.path.join(ROOT, "profiles", "pmc_traffic_tg256.md")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--precision", choices=["both", "fp64", "mixed"], default="both")
    ap.add_argument("--profile-kernel", default="k_dir_apply")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=128, help="grid of the bounded CPU baseline sample")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--adjoint", choices=["c4", "c5", "none"], default="c4",
                    help="adjoint workload timed beside the forward step (c5: 512^3, about 160 GB of HBM)")
    ap.add_argument("--adjoint-fd", type=int, default=0,
                    help="central finite-difference checks of the adjoint gradient along this many seeded directions")
    ap.add_argument("--adjoint-tl", type=int, default=0,
                    help="adjoint-versus-tangent-linear checks along this many seeded directions")
    ap.add_argument("--adjoint-only", action="store_true", help="skip the forward benchmark")
    ap.add_argument("--predictor-cg-steps", type=int, default=20,
                    help="FP64 steps timed with the CG Helmholtz predictor beside the line solves (0: skip)")
    return ap.parse_args()


ADJOINT_SEED = 20261203
# 128^3 Re=400, 200 steps, checkpoint every 20, FP64, gradient of the final kinetic energy
# w.r.t. the initial velocity; 512^3 Re=1600, 100 steps, checkpoint every 10, mixed precision,
# gradient w.r.t. a seeded forcing field f = (A sin(kx x) sin(ky y), 0, 0), A = 1e-2.
ADJOINT_CASES = {
    "c4": dict(n=128, nu=1.0 / 400.0, steps=200, every=20, mixed=False, forced=False, tol=1e-12),
    "c5": dict(n=512, nu=1.0 / 1600.0, steps=100, every=10, mixed=True, forced=True, tol=1e-10),
}


def run_adjoint(sa, ctx, name, fd_dirs, tl_dirs=0):
    """Forward run with checkpoints, then the reverse sweep (window recompute + adjoint steps).
    Initial state: the Taylor-Green field plus the seeded divergence-free perturbation of RMS
    1e-2. Optional central finite differences of the final kinetic energy along seeded
    directions (initial velocity, or forcing for the forced case)."""
    import torch

    from solver_amd import inputs

    cs = ADJOINT_CASES[name]
    n, nu, steps, every = cs["n"], cs["nu"], cs["steps"], cs["every"]
    prec = sa.MIXED if cs["mixed"] else sa.FP64
    L = 2 * math.pi
    dt, ns = sa.taylor_green_dt(n, n, n, 0.5, steps * (0.5 / (3 * n / L)) * (1 - 1e-14))
    assert ns == steps, ns
    st = sa.Stepper(ctx, sa.TAYLOR_GREEN, n, n, n, nu, dt, 400, cs["tol"], prec)
    st.initialize(convention=1)
    vel, p, _, _, _ = st.download()
    st.close()
    pert = inputs.perturbation(n, ADJOINT_SEED, 1e-2)
    u0 = [torch.from_numpy(a + b).cuda() for a, b in zip(vel, pert)]
    p0 = torch.from_numpy(p).cuda()
    del vel, pert, p
    adj = sa.Adjoint(ctx, n, n, n, nu, dt, every, 400, cs["tol"], prec)
    if cs["forced"]:
        adj.set_forcing([torch.from_numpy(f).cuda() for f in inputs.forcing(n, ADJOINT_SEED, 1e-2)])
    torch.cuda.synchronize()
    print(f"# adjoint {name}: inputs ready, forward pass of {steps} steps", flush=True)
    t0 = time.perf_counter()
    ke, m = adj.forward(u0, p0, steps)
    t1 = time.perf_counter()
    print(f"# adjoint {name}: forward {t1 - t0:.2f} s, reverse sweep", flush=True)
    gu, _, gf = adj.backward(want_pressure=False, want_forcing=cs["forced"])
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    out = {"workload": f"{n}^3 Re={round(1 / nu)} {steps} steps, checkpoint every {every}, "
                       f"{'mixed' if cs['mixed'] else 'fp64'}, d KE(T) / d {'f' if cs['forced'] else 'u0'}",
           "n": n, "steps": steps, "checkpoint_every": every, "dt": dt, "poisson_tolerance": cs["tol"],
           "forward_s": t1 - t0, "reverse_s": t2 - t1, "reverse_s_per_step": (t2 - t1) / steps,
           "gradient_s_per_step": (t2 - t0) / steps, "final_kinetic_energy": ke,
           "max_div_linf": m.max_divergence_linf}
    if cs["forced"]:
        out["forcing_wavenumbers"] = list(inputs.forcing_wavenumbers(ADJOINT_SEED))
    grad = gf if cs["forced"] else gu

    def direction(d):
        # velocity directions: seeded divergence-free fields of RMS 1; forcing directions: seeded
        # forcing patterns of unit amplitude
        return inputs.forcing(n, ADJOINT_SEED + d, 1.0) if cs["forced"] else \
            inputs.perturbation(n, ADJOINT_SEED + d, 1.0)

    errs = []
    for d in range(1, fd_dirs + 1):
        dirn = [torch.from_numpy(a).cuda() for a in direction(d)]
        pred = float(sum((g * x).sum() for g, x in zip(grad, dirn)))
        best = None
        for eps in (1e-5, 1e-6, 1e-7):
            ks = []
            for sgn in (1.0, -1.0):
                if cs["forced"]:
                    base = inputs.forcing(n, ADJOINT_SEED, 1e-2)
                    adj.set_forcing([torch.from_numpy(b).cuda() + sgn * eps * x for b, x in zip(base, dirn)])
                    ks.append(adj.forward(u0, p0, steps)[0])
                else:
                    ks.append(adj.forward([a + sgn * eps * x for a, x in zip(u0, dirn)], p0, steps)[0])
            e = abs((ks[0] - ks[1]) / (2 * eps) - pred) / abs(pred)
            best = e if best is None else min(best, e)
        errs.append(best)
        print(f"# adjoint {name}: direction {d}: <grad, d> = {pred:.10e}, best central-difference rel. error "
              f"{best:.2e}", flush=True)
    if errs:
        out["fd_directions"] = len(errs)
        out["fd_max_rel_error"] = max(errs)
        out["fd_eps"] = [1e-5, 1e-6, 1e-7]
        # back to the base trajectory for the tangent checks
        if cs["forced"]:
            adj.set_forcing([torch.from_numpy(f).cuda() for f in inputs.forcing(n, ADJOINT_SEED, 1e-2)])
        adj.forward(u0, p0, steps)
    # tangent-linear consistency: <grad, d> against the device tangent along d
    tl_errs = []
    for d in range(1, tl_dirs + 1):
        dirn = [torch.from_numpy(a).cuda() for a in direction(d)]
        pred = float(sum((g * x).sum() for g, x in zip(grad, dirn)))
        t3 = time.perf_counter()
        _, dke = adj.tangent(dforce=dirn) if cs["forced"] else adj.tangent(dvel=dirn)
        t4 = time.perf_counter()
        tl_errs.append(abs(dke - pred) / abs(pred))
        out["tangent_s"] = t4 - t3
        print(f"# adjoint {name}: direction {d}: <grad, d> = {pred:.10e}, tangent {dke:.10e}, "
              f"rel. difference {tl_errs[-1]:.2e}", flush=True)
    if tl_errs:
        out["tl_directions"] = len(tl_errs)
        out["tl_max_rel_difference"] = max(tl_errs)
    adj.close()
    del u0, p0, grad, gu, gf
    torch.cuda.empty_cache()
    return out


# Algorithmic HBM bytes per launch of the finest-level kernels: (distinct arrays read + arrays
# written) x cells x element size; halo re-reads and caches excluded. t = multigrid element size.
def kernel_bytes(name, n_cells, mixed):
    f64, t = 8, (4 if mixed else 8)
    table = {
        "k_dir_apply": t + 5 * f64,                     # z, d_old, p read; d, p, q written
        "k_apply_dot": 2 * f64,                         # d read; q written
        "k_direction": t + 2 * f64 + 2 * f64,           # z, d_old, p read; d_new, p written
        "k_update": 3 * f64 + (4 if mixed else 0),      # r, q read; r (+ FP32 copy) written
        "k_down_fused": 2 * t + t / 8,                  # b read; x2, coarse b written
        "k_up_fused": 3 * t + t / 8,                    # b, x2, coarse x read; x5 written
        "k_smooth12": 2 * t,                            # b read; x2 written (unfused levels)
        "k_residual_restrict": 2 * t + t / 8,           # b, x2 read; coarse b written
        "k_prolong_smooth": 3 * t + t / 8,              # b, x2, coarse x read; x4 written
        "k_smooth_last": 3 * t,                         # b, x4 read; x5 written
        "k_predictor_rhs": 7 * f64,                     # q, 2 other components, p, A_prev; A, rhs
        "k_line_solve_x": 3 * 2 * f64,                  # 3 components: line read + written once
        "k_line_solve_s": 3 * 2 * f64,
    }
    return table[name] * n_cells


def step_bytes_model(n, iters, mixed):
    """Algorithmic HBM bytes of one step (passes over an n^3 FP64 array):
    predictor RHS 3 x 7, ADI staging copies 3 x 2 + sweeps 9 x 2, pressure RHS 4, Poisson
    setup/finish 12 (+1.5 FP32 cast), correction 9, pressure update 4, diagnostics 6; per MGPCG
    iteration: fused direction/operator 6 (z, d_old, p read; d, p, q written; z in FP32 and an
    FP32 copy of r in mixed: 5.5 + 3.5) and residual update 3, and the V-cycle 10.25 x 8/7
    (hierarchy) in the multigrid precision (x 1/2 for FP32)."""
    fixed = 21 + 6 + 18 + 4 + 12 + (1.5 if mixed else 0.0) + 9 + 4 + 6
    per_it = 9.0 + 10.25 * 8.0 / 7.0 * (0.5 if mixed else 1.0)
    return (fixed + iters * per_it) * n ** 3 * 8


def measured_traffic(kernel, mixed):
    """HBM bytes per finest-grid launch of `kernel` from the committed rocprofv3 counter summary
    (rows are named with template arguments, e.g. k_dir_apply<float, 8, true>; the first row of
    a kernel is its finest-grid launch)."""
    try:
        text = open(TRAFFIC_TABLE).read()
    except OSError:
        return None
    elem = "float" if mixed else "double"
    for line in text.splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) < 6:
            continue
        name = cells[0]
        if name == kernel or name.startswith(f"{kernel}<{elem}") or (name.startswith(f"{kernel}<") and "," not in name
                                                                       and elem in name):
            try:
                return (float(cells[4]) + float(cells[5])) * 1e6
            except ValueError:
                return None
    return None


def host_cpu():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(n_target, n, nu, cfl):
    """The reference's own CPU path (oracle/_ref/ref_golden, built from the reference sources,
    -O3, no fast-math) on this host, one thread, for one step of the same case at n^3; the
    steps/s is scaled to n_target^3 by the cell ratio. Falls back to the oracle restatement."""
    from oracle import refrun

    L = 2 * math.pi
    final_time = 100 * (cfl / (3 * n / L)) * (1 - 1e-14)
    scale = (n / n_target) ** 3
    if refrun.available():
        t0 = time.perf_counter()
        _, meta = refrun.run("tg_steps", (), (), timeout=1800, nx=n, ny=n, nz=n, viscosity=nu, cfl=cfl,
                             final_time=final_time, tol=1e-10, maxit=400, steps=1)
        wall = time.perf_counter() - t0
        el = float(meta.get("elapsed_seconds", wall))
        kind, what = "reference", "the reference CPU sources built -O3 without fast-math"
        iters = meta.get("pressure_iterations")
    else:
        from oracle import ofref

        c = ofref.tg_config(n, n, n, nu, cfl, final_time, 400, 1e-10)
        s = ofref.tg_init(c, convention=1)
        t0 = time.perf_counter()
        ofref.tg_steps(c, s, 1)
        el = time.perf_counter() - t0
        kind, what = "port", "the oracle's FP64 restatement of the reference"
        iters = s.metrics.pressure_iterations
    return {"value": scale / el, "unit": "steps/s", "cores": 1, "kind": kind,
            "sample": f"1 step of the same Taylor-Green case at {n}^3 ({n ** 3} cells, FP64, MGPCG tol 1e-10, "
                      f"{iters} iterations) with {what}, 1 thread: {el:.2f} s; value = 1/{el:.2f} s scaled to "
                      f"{n_target}^3 by the cell ratio {scale:g}",
            "host_cpu": host_cpu(), "host_logical_cpus": os.cpu_count()}


def run_precision(sa, ctx, n, nu, dt, precision, warmup, steps, prof_kernel, predictor_cg=False):
    import torch

    st = sa.Stepper(ctx, sa.TAYLOR_GREEN, n, n, n, nu, dt, 400, 1e-10, precision)
    if predictor_cg:  # the Metal backend's predictor at its FP64 tolerance
        st.set_predictor(sa.PREDICTOR_CG, 1e-12, 80)
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


def copy_calibration():
    import torch

    n = 1 << 27  # 2^27 doubles = 1 GiB
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8, "copies": 3}))
    return 0


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU: the path is measured on one MI355X "
                                   "(multi-GPU sharding is a design note, not a measured configuration)",
                          "n_gpus": a.gpus}))
        return 2
    if a.copy_calibration:
        return copy_calibration()
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
    if a.adjoint_only:
        r = run_adjoint(sa, ctx, a.adjoint if a.adjoint != "none" else "c4", a.adjoint_fd, a.adjoint_tl)
        print(json.dumps({"metric": "adjoint s/step", "value": r["reverse_s_per_step"], "unit": "s/step",
                          "higher_is_better": False, "adjoint": r}))
        return 0
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

    def step_hbm(rr, mixed):
        it = sum(rr["iters"]) / len(rr["iters"])
        b = step_bytes_model(n, it, mixed)
        gbs = b / (rr["seconds"] / a.steps) / 1e9
        return {"algorithmic_bytes_per_step": b, "achieved_GBs": gbs, "peak_GBs": HBM_PEAK_GBS,
                "frac": gbs / HBM_PEAK_GBS}

    prof = r["prof"]
    avg_ms = prof["total_ms"] / max(prof["launches"], 1)
    nbytes = kernel_bytes(a.profile_kernel, n_cells, main_prec == sa.MIXED)
    achieved = nbytes / (avg_ms * 1e-3) / 1e9 if prof["launches"] else None
    out = {
        "metric": METRIC,
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
        "step_hbm": step_hbm(r, main_prec == sa.MIXED),
        "roofline": {"bound": "hbm", "kernel": a.profile_kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": measured_traffic(a.profile_kernel, main_prec == sa.MIXED),
                     "algorithmic_bytes_per_launch": nbytes, "avg_launch_ms": avg_ms,
                     "launches": prof["launches"]},
        "adjoint_s_per_step": None,
    }
    if sa.FP64 in res:
        f = res[sa.FP64]
        out["fp64"] = {"steps_per_s": a.steps / f["seconds"], "ms_per_step": 1e3 * f["seconds"] / a.steps,
                       "mgpcg_iterations_per_step": sum(f["iters"]) / len(f["iters"]), "max_div_linf": f["div_linf"],
                       "step_hbm": step_hbm(f, False)}
        if sa.MIXED in res:
            out["mixed_speedup_vs_fp64"] = value / out["fp64"]["steps_per_s"]
            out["mixed_vs_fp64_rel_l2"] = {k: rel_l2(x, y) for k, x, y in
                                           zip("uvwp", res[sa.MIXED]["fields"], res[sa.FP64]["fields"])}
    if a.predictor_cg_steps > 0 and sa.FP64 in res:
        # the CG Helmholtz predictor in place of the line solves, FP64, on the same first steps
        k = min(a.predictor_cg_steps, a.steps)
        rc = run_precision(sa, ctx, n, nu, dt, sa.FP64, a.warmup, k, "", predictor_cg=True)
        rf = run_precision(sa, ctx, n, nu, dt, sa.FP64, a.warmup, k, "")
        out["fp64_cg_predictor"] = {
            "steps": k, "steps_per_s": k / rc["seconds"], "ms_per_step": 1e3 * rc["seconds"] / k,
            "adi_ms_per_step": 1e3 * rf["seconds"] / k, "max_div_linf": rc["div_linf"],
            "vs_adi_rel_l2": {key: rel_l2(x, y) for key, x, y in zip("uvwp", rc["fields"], rf["fields"])}}
        print(f"# fp64 with the CG predictor: {k / rc['seconds']:.2f} steps/s ({1e3 * rc['seconds'] / k:.2f} ms/step "
              f"vs {1e3 * rf['seconds'] / k:.2f} with the line solves)", flush=True)
    if a.adjoint != "none":
        for k in list(res):
            res[k].pop("fields", None)
        r_adj = run_adjoint(sa, ctx, a.adjoint, a.adjoint_fd, a.adjoint_tl)
        print(f"# adjoint {a.adjoint}: forward {r_adj['forward_s']:.2f} s, reverse {r_adj['reverse_s']:.2f} s "
              f"({r_adj['reverse_s_per_step'] * 1e3:.1f} ms/step)", flush=True)
        out["adjoint_s_per_step"] = r_adj["reverse_s_per_step"]
        out["adjoint"] = r_adj
    if not a.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(n, a.cpu_n, nu, cfl)
        except Exception as e:  # reported, never fatal to the GPU measurement
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
