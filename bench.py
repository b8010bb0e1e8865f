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
import re
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "3D TG 256³ steps/s mixed vs FP64 at ‖div u‖∞≤1e-10; HBM GB/s; adjoint s/step"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
PROFILE_STRIDE = 8  # live kernel timing samples every 8th launch of the kernel
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
    ap.add_argument("--cpu-steps", type=int, default=1,
                    help="steps of the benchmark case the reference CPU path is timed for (one is about 50 s)")
    ap.add_argument("--no-c1", action="store_true", help="skip the 2D 128^2 (C1) run")
    ap.add_argument("--fp64-chebyshev-steps", type=int, default=20,
                    help="FP64 steps timed with the mixed mode's Chebyshev sweep weights (0: skip)")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--adjoint", choices=["c4", "c5", "none"], default="c4",
                    help="adjoint workload timed beside the forward step (c5: 512^3, about 160 GB of HBM)")
    ap.add_argument("--adjoint-fd", type=int, default=0,
                    help="central finite-difference checks of the adjoint gradient along this many seeded directions")
    ap.add_argument("--adjoint-tl", type=int, default=0,
                    help="adjoint-versus-tangent-linear checks along this many seeded directions")
    ap.add_argument("--adjoint-only", action="store_true", help="skip the forward benchmark")
    ap.add_argument("--adjoint-steps", type=int, default=0,
                    help="steps of the adjoint workload instead of its own count (short profiling runs)")
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


def run_adjoint(sa, ctx, name, fd_dirs, tl_dirs=0, steps_override=0):
    """Forward run with checkpoints, then the reverse sweep (window recompute + adjoint steps).
    Initial state: the Taylor-Green field plus the seeded divergence-free perturbation of RMS
    1e-2. Optional central finite differences of the final kinetic energy along seeded
    directions (initial velocity, or forcing for the forced case)."""
    import torch

    from solver_amd import inputs

    cs = ADJOINT_CASES[name]
    n, nu, steps, every = cs["n"], cs["nu"], cs["steps"], cs["every"]
    if steps_override:
        steps = steps_override
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
        "k_dir_apply": t + 2 * f64,                     # z, d_old read; d written
        "k_apply_dot": 2 * f64,                         # d read; q written
        "k_direction": t + 2 * f64,                     # z, d_old read; d written
        "k_update": 3 * f64 + (4 if mixed else 0),      # r, d read; r (+ FP32 copy) written
        "k_down_fused": 2 * t + t / 8,                  # b read; x2, coarse b written
        "k_up_fused": 3 * t + t / 8,                    # b, x2, coarse x read; x5 written
        "k_smooth12": 2 * t,                            # b read; x2 written (unfused levels)
        "k_residual_restrict": 2 * t + t / 8,           # b, x2 read; coarse b written
        "k_prolong_smooth": 3 * t + t / 8,              # b, x2, coarse x read; x4 written
        "k_smooth_last": 3 * t,                         # b, x4 read; x5 written
        "k_predictor_rhs": 7 * f64,                     # q, 2 other components, p, A_prev; A, rhs
        "k_line_solve_x": 3 * 2 * f64,                  # 3 components: line read + written once
        "k_line_solve_s": 3 * 2 * f64,
        "k_finish_periodic": 13 * f64,                  # pred x3, phi, prhs, old u x3, p read; u x3, p written
        "k_helm_conv_x": 2 * f64,                       # one component: line read + written once
        "k_helm_conv_s": 2 * f64,
    }
    return table[name] * n_cells


def step_bytes_model(n, iters, mixed):
    """Algorithmic HBM bytes of one step by the stage model of the path (passes over an n^3
    FP64 array, each stage counted as its distinct arrays read + written once):
    predictor RHS with the three components fused 13 (u, v, w, p, A_prev x 3 read; A x 3 and
    rhs x 3 written); ADI 6 per component (rhs read, line written, staged once) = 18; pressure
    RHS 4 + its mean 1; correction with max|du| and the divergence 13; pressure update 4;
    diagnostics 6. Per MGPCG iteration: CG vector work 11 and the V-cycle 15.1 (13.25 on the
    finest level x 8/7 for the hierarchy), the V-cycle halved when it runs in FP32."""
    fixed = 13 + 18 + 5 + 13 + 4 + 6
    per_it = 11.0 + 15.1 * (0.5 if mixed else 1.0)
    return (fixed + iters * per_it) * n ** 3 * 8


def step_kernel_bytes(n, iters, mixed):
    """HBM bytes the launches of one step read and write by their own arrays (each kernel's
    distinct arrays read + written once, the fused kernels counted as fused): predictor RHS
    3 x 7, periodic line sweeps 9 x 2, pressure RHS 4, Poisson set-up 4 (+1.5 FP32 cast), the
    pressure pass's write 1 and mean removal 2, correction with the pressure update and the
    diagnostics in one pass 13;
    per MGPCG iteration the fused direction/operator 3 (2.5 mixed), the residual update with the
    operator recomputed 3 (3.5 mixed), the pressure pass's read of the iteration's direction 1,
    and the fused V-cycle levels 2.125 (down) + 3.125 (up) x 8/7 in the multigrid precision."""
    fixed = 21 + 18 + 4 + 4 + (1.5 if mixed else 0.0) + 1 + 2 + 13
    per_it = 7.0 + 5.25 * 8.0 / 7.0 * (0.5 if mixed else 1.0)
    return (fixed + iters * per_it) * n ** 3 * 8


def measured_traffic(kernel, mixed):
    """HBM bytes per finest-grid launch of `kernel` from the committed rocprofv3 counter summary
    (rows are named with template arguments, e.g. k_dir_apply_lds<float>; the variants of one
    operation (_lds, _zw) answer for its name; the first row of a kernel is its finest-grid
    launch)."""
    try:
        text = open(TRAFFIC_TABLE).read()
    except OSError:
        return None
    elem = "float" if mixed else "double"
    for line in text.splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) < 6:
            continue
        name = re.sub(r"_(lds|zw)(?=<|$| )", "", cells[0])
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


def cpu_baseline(n, nu, cfl, steps):
    """The reference's own CPU path (oracle/_ref/ref_golden, built from the reference sources,
    -O3, no fast-math) on this host, one thread, for `steps` steps of the benchmark case itself
    (n^3, the same dt, tolerance and initial field convention as the GPU run). Falls back to
    the oracle restatement where the reference build is absent."""
    from oracle import refrun

    L = 2 * math.pi
    final_time = 100 * (cfl / (3 * n / L)) * (1 - 1e-14)
    if refrun.available():
        print(f"# cpu baseline: the reference CPU path, {steps} step(s) at {n}^3", flush=True)
        t0 = time.perf_counter()
        _, meta = refrun.run("tg_steps", (), (), timeout=1800, nx=n, ny=n, nz=n, viscosity=nu, cfl=cfl,
                             final_time=final_time, tol=1e-10, maxit=400, steps=steps)
        wall = time.perf_counter() - t0
        el = float(meta.get("elapsed_seconds", wall))
        kind, what = "reference", "the reference CPU sources built -O3 without fast-math"
        iters = meta.get("pressure_iterations")
    else:
        from oracle import ofref

        c = ofref.tg_config(n, n, n, nu, cfl, final_time, 400, 1e-10)
        s = ofref.tg_init(c, convention=0)
        t0 = time.perf_counter()
        ofref.tg_steps(c, s, steps)
        el = time.perf_counter() - t0
        kind, what = "port", "the oracle's FP64 restatement of the reference"
        iters = s.metrics.pressure_iterations
    return {"value": steps / el, "unit": "steps/s", "cores": 1, "kind": kind,
            "sample": f"{steps} step(s) of the benchmark case itself ({n}^3 = {n ** 3} cells, FP64, MGPCG tol "
                      f"1e-10, {iters} iterations in the last step, the reference's own initial field) with "
                      f"{what}, 1 thread: {el:.2f} s",
            "seconds": el, "host_cpu": host_cpu(), "host_logical_cpus": os.cpu_count()}


C1 = dict(n=128, nu=0.01, cfl=0.5, steps=200, tol=1e-10, maxit=200)


def c1_run(sa, ctx, with_cpu):
    """The 2D Taylor-Green benchmark configuration (128^2, viscosity 0.01, CFL 0.5, 200 steps,
    MGPCG tolerance 1e-10, shifted initial field): the whole run on the device in FP64, and the
    reference CPU path over the same 200 steps on this host (one thread)."""
    import torch

    n, steps = C1["n"], C1["steps"]
    L = 2 * math.pi
    final_time = steps * (C1["cfl"] / (2 * n / L)) * (1 - 1e-12)
    dt, ns = sa.taylor_green_dt(n, n, 1, C1["cfl"], final_time)
    assert ns == steps
    st = sa.Stepper(ctx, sa.TAYLOR_GREEN, n, n, 1, C1["nu"], dt, C1["maxit"], C1["tol"], sa.FP64)
    st.initialize(convention=1)
    st.run(1)  # warm-up of the kernels' first launches
    st.initialize(convention=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m = st.run(steps)
    ctx.sync()
    gpu_s = time.perf_counter() - t0
    st.close()
    out = {"workload": "2D periodic Taylor-Green 128^2, viscosity 0.01, CFL 0.5, 200 steps, MGPCG tol 1e-10, fp64",
           "gpu_seconds": gpu_s, "gpu_steps_per_s": steps / gpu_s, "max_div_linf": m.max_divergence_linf}
    if with_cpu:
        from oracle import refrun, ofref

        if refrun.available():
            c = ofref.tg_config(n, n, 1, C1["nu"], C1["cfl"], final_time, C1["maxit"], C1["tol"])
            s0 = ofref.tg_init(c, convention=1)
            shp = [ofref.storage_shape(n, n, 1, loc) for loc in (1, 2, 3, 0, 1, 2, 3)]
            print("# cpu baseline: the reference CPU path, C1 200 steps", flush=True)
            _, meta = refrun.run("tg_steps", s0.vel + [s0.p] + s0.adv, shp, timeout=1800, nx=n, ny=n, nz=1,
                                 viscosity=C1["nu"], cfl=C1["cfl"], final_time=final_time, tol=C1["tol"],
                                 maxit=C1["maxit"], steps=steps, has_prev=0)
            el = float(meta["elapsed_seconds"])
            out["cpu_reference_seconds"] = el
            out["cpu_reference_steps_per_s"] = steps / el
            out["gpu_vs_cpu"] = el / gpu_s
    return out


def run_precision(sa, ctx, n, nu, dt, precision, warmup, steps, prof_kernel, predictor_cg=False, chebyshev=False):
    import torch

    if chebyshev:
        os.environ["SOLVER_AMD_FP64_CHEBYSHEV"] = "1"
    try:
        st = sa.Stepper(ctx, sa.TAYLOR_GREEN, n, n, n, nu, dt, 400, 1e-10, precision)
    finally:
        os.environ.pop("SOLVER_AMD_FP64_CHEBYSHEV", None)
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
    # every 8th launch is bracketed with events: an event record idles the device for a few
    # microseconds, which timing every launch would add to the steps being measured
    ctx.profile_stride(PROFILE_STRIDE)
    iters = []
    t0 = time.perf_counter()
    for _ in range(steps):
        m = st.run(1)
        iters.append(m.pressure_iterations)
        div_inf = max(div_inf, m.divergence_linf)
    ctx.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    durs = ctx.profile_durations()
    ctx.profile_kernel("")
    vel, p, _, _, _ = st.download()
    st.close()
    return {"seconds": t1 - t0, "iters": iters, "div_linf": div_inf, "div_l2": m.divergence_l2,
            "durations": durs, "fields": vel + [p], "max_cfl": m.max_cfl}


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
        r = run_adjoint(sa, ctx, a.adjoint if a.adjoint != "none" else "c4", a.adjoint_fd, a.adjoint_tl,
                        a.adjoint_steps)
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
        kb = step_kernel_bytes(n, it, mixed)
        sec = rr["seconds"] / a.steps
        gbs = b / sec / 1e9
        return {"algorithmic_bytes_per_step": b, "achieved_GBs": gbs, "peak_GBs": HBM_PEAK_GBS,
                "frac": gbs / HBM_PEAK_GBS, "model": "stage passes (fused predictor, 6-pass ADI, 11 + 15.1 per "
                "MGPCG iteration)", "kernel_bytes_per_step": kb, "kernel_bytes_GBs": kb / sec / 1e9,
                "kernel_bytes_frac": kb / sec / 1e9 / HBM_PEAK_GBS}

    # sampled launches (every PROFILE_STRIDE-th) that did work: after convergence the host has
    # already queued one more iteration whose kernels return at entry (a few microseconds);
    # they are not the kernel's duration and are left out of the average
    durs = r["durations"]
    med = sorted(durs)[len(durs) // 2] if durs else 0.0
    work = [d for d in durs if d > 0.25 * med]
    avg_ms = sum(work) / len(work) if work else 0.0
    nbytes = kernel_bytes(a.profile_kernel, n_cells, main_prec == sa.MIXED)
    achieved = nbytes / (avg_ms * 1e-3) / 1e9 if work else None
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
                     "algorithmic_bytes_per_launch": nbytes, "avg_launch_ms": avg_ms, "median_launch_ms": med,
                     "launches": len(work), "early_exit_launches": len(durs) - len(work)},
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
    if a.fp64_chebyshev_steps > 0 and sa.FP64 in res and sa.MIXED in res:
        # like for like: FP64 with the mixed mode's sweep weights separates the two sources of
        # the mixed speed-up (smoother weights vs FP32 arithmetic); not the reference's smoother
        k = min(a.fp64_chebyshev_steps, a.steps)
        rch = run_precision(sa, ctx, n, nu, dt, sa.FP64, a.warmup, k, "", chebyshev=True)
        rf = run_precision(sa, ctx, n, nu, dt, sa.FP64, a.warmup, k, "")
        rm = run_precision(sa, ctx, n, nu, dt, sa.MIXED, a.warmup, k, "")
        cheb = k / rch["seconds"]
        out["fp64_chebyshev_weights"] = {
            "steps": k, "steps_per_s": cheb, "ms_per_step": 1e3 * rch["seconds"] / k,
            "mgpcg_iterations_per_step": sum(rch["iters"]) / k, "max_div_linf": rch["div_linf"],
            "fp64_reference_weights_steps_per_s": k / rf["seconds"], "mixed_steps_per_s": k / rm["seconds"],
            "speedup_from_weights": cheb / (k / rf["seconds"]),
            "speedup_from_fp32": (k / rm["seconds"]) / cheb,
            "vs_fp64_rel_l2": {key: rel_l2(x, y) for key, x, y in zip("uvwp", rch["fields"], rf["fields"])}}
        print(f"# fp64 with Chebyshev weights: {cheb:.2f} steps/s, {sum(rch['iters']) / k:.1f} iterations/step; "
              f"speed-up from the weights {out['fp64_chebyshev_weights']['speedup_from_weights']:.2f}x, "
              f"from FP32 {out['fp64_chebyshev_weights']['speedup_from_fp32']:.2f}x", flush=True)
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
        r_adj = run_adjoint(sa, ctx, a.adjoint, a.adjoint_fd, a.adjoint_tl, a.adjoint_steps)
        print(f"# adjoint {a.adjoint}: forward {r_adj['forward_s']:.2f} s, reverse {r_adj['reverse_s']:.2f} s "
              f"({r_adj['reverse_s_per_step'] * 1e3:.1f} ms/step)", flush=True)
        out["adjoint_s_per_step"] = r_adj["reverse_s_per_step"]
        out["adjoint"] = r_adj
    if not a.no_c1:
        out["c1"] = c1_run(sa, ctx, with_cpu=not a.no_cpu_baseline)
        print(f"# c1 (2D 128^2, 200 steps): {out['c1']['gpu_steps_per_s']:.0f} steps/s on the device", flush=True)
    if not a.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(n, nu, cfl, a.cpu_steps)
            out["cpu_baseline"]["gpu_vs_cpu"] = value / out["cpu_baseline"]["value"]
        except Exception as e:  # reported, never fatal to the GPU measurement
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
