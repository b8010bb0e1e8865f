"""The small benchmark configurations run in full on the device, against the reference's own
runs (golden fixtures written by the reference CPU sources) and the oracle.

* C1: 2D periodic Taylor-Green 128^2 on [0, 2 pi]^2, u = sin x cos y, v = -cos x sin y, viscosity
  0.01, CFL 0.5, 200 steps, TVD advection, MGPCG tolerance 1e-10 (fixture tg2d_128_c1): fields
  within 1e-10 relative L2 of the reference after 200 steps, and the oracle's MGPCG iteration
  count at every step.
* C2: lid-driven cavity 128^2, Re 100, run to steady state by the reference's rule (step >=
  min_steps and max|du| <= tolerance) at max|du| <= 1e-8 with Poisson tolerance 1e-10
  (fixture cavity_c2_128), and the reference's benchmark setting 1e-7 / 1e-9 through the
  solver_cavity tool with the Re 100 centreline envelope check (fixture cavity_bench_128).
"""
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest

import solver_amd as sa
from oracle import ofref
from tests.golden import fixtures as fx
from tests.gpu_util import rel_l2_active

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "solver_amd", "cpp", "build", "bin")


@pytest.fixture(scope="module")
def ctx():
    c = sa.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("precision", [sa.FP64, sa.MIXED])
def test_c1_taylor_green_2d_128_full_run(ctx, precision):
    f = fx.load("tg2d_128_c1")
    p = f.params
    dt, steps = sa.taylor_green_dt(128, 128, 1, p["cfl"], p["final_time"])
    assert steps == 200 and dt == pytest.approx(f.meta["dt"], rel=1e-15)
    st = sa.Stepper(ctx, sa.TAYLOR_GREEN, 128, 128, 1, p["viscosity"], dt, p["maxit"], p["tol"], precision)
    st.initialize(1)
    vel0, p0, _, _, _ = st.download()
    # the device's shifted initial field is the fixture's input
    for a, k in zip([vel0[0], vel0[1], p0], "uvp"):
        assert np.max(np.abs(a - f.inp[k])) <= 1e-15, k
    iters = []
    for _ in range(steps):
        m = st.run(1)
        iters.append(m.pressure_iterations)
        assert m.divergence_linf <= 1e-10
    vel, pt, _, _, _ = st.download()
    st.close()
    assert m.step == f.meta["step"] == 200
    if precision == sa.FP64:
        assert m.pressure_iterations == f.meta["pressure_iterations"]
        c = ofref.tg_config(128, 128, 1, p["viscosity"], p["cfl"], p["final_time"], p["maxit"], p["tol"])
        s = ofref.tg_init(c, convention=1)
        ref_iters = []
        for _ in range(steps):
            ofref.tg_steps(c, s, 1)
            ref_iters.append(s.metrics.pressure_iterations)
        assert iters == ref_iters
    for a, k in zip([vel[0], vel[1], pt], "uvp"):
        err = rel_l2_active(a, f.out[k], mean_removed=(k == "p"))
        assert err <= 1e-10, f"{k}: rel L2 {err:.3e}"
    # energy decay against the analytic solution exp(-4 nu t) (solver/taylor_green.cpp:258-267)
    L = 2 * math.pi
    ke = 0.5 * (np.mean(vel[0][1:-1, 1:-1, 1:-2] ** 2) + np.mean(vel[1][1:-1, 1:-2, 1:-1] ** 2)) * L * L
    ke0 = 0.5 * (np.mean(vel0[0][1:-1, 1:-1, 1:-2] ** 2) + np.mean(vel0[1][1:-1, 1:-2, 1:-1] ** 2)) * L * L
    assert abs(ke / ke0 - math.exp(-4 * p["viscosity"] * m.time)) <= 5e-3


def cavity_steps_until(ctx, p, precision=sa.FP64):
    """Device stepper run by run_lid_driven_cavity's rule: stop at step >= min_steps with
    max|du| <= steady_tolerance, or at max_steps; returns (stepper, metrics, error or None)."""
    dt = sa.lid_driven_cavity_dt(p["nx"], p["ny"], p["lid"], p["cfl"])
    nu = p["lid"] / p["reynolds"]
    st = sa.Stepper(ctx, sa.LID_DRIVEN_CAVITY, p["nx"], p["ny"], 1, nu, dt, p["maxit"], p["tol"], precision,
                    lid_velocity=p["lid"])
    st.initialize(0)
    m = None
    while True:
        try:
            m = st.run(1)
        except sa.SolverError as e:
            return st, m, str(e)
        if (m.step >= p["min_steps"] and m.max_velocity_change <= p["steady_tolerance"]) or m.step >= p["max_steps"]:
            return st, m, None


def test_c2_cavity_128_to_steady_state(ctx):
    """The 128^2 lid-driven cavity at Re 100 run to steady state at max|du| <= 1e-8 with Poisson
    tolerance 1e-10 (the C2 configuration): the device follows the
    reference step for step and ends where the reference ends -- the reference's own run stops
    before reaching 1e-8, when the pressure solve meets its 1e-30 breakdown guard above the
    tolerance, with its non-convergence error; the device raises the same error at the same
    step and leaves the same fields."""
    f = fx.load("cavity_c2_128")
    p = f.params
    print(f"C2 cavity: reference outcome step {f.meta['step']}, error {f.meta.get('error')}", flush=True)
    st, m, err = cavity_steps_until(ctx, p)
    vel, pt, _, _, mm = st.download()
    st.close()
    if "error" in f.meta:
        assert err is not None, f"the device reached step {mm.step} without the reference's error"
        assert err.split(", relative_residual")[0] == f.meta["error"].split(", relative_residual")[0]
    else:
        assert err is None, err
        assert bool(f.meta["steady"]) == (mm.max_velocity_change <= p["steady_tolerance"])
    assert mm.step == f.meta["step"]
    assert mm.pressure_iterations == f.meta["pressure_iterations"]
    assert abs(mm.max_velocity_change - f.meta["max_velocity_change"]) <= 1e-6 * f.meta["max_velocity_change"]
    for a, k in zip([vel[0], vel[1], pt], "uvp"):
        err_k = rel_l2_active(a, f.out[k], mean_removed=(k == "p"))
        assert err_k <= 1e-9, f"{k}: rel L2 {err_k:.3e}"


def test_c2_cavity_tool_reference_benchmark_setting():
    """The reference's own benchmark configuration (benchmarks/lid_driven_cavity_re100_128.cfg:
    steady at max|du| <= 1e-7, Poisson tolerance 1e-9, min_steps 400) through the solver_cavity
    tool: the steady step, the centreline samples and the Re 100 envelope verdict of the
    reference's run."""
    f = fx.load("cavity_bench_128")
    p = f.params
    with tempfile.TemporaryDirectory() as td:
        cfg = os.path.join(td, "cavity.cfg")
        with open(cfg, "w") as fh:
            fh.write(f"nx = {p['nx']}\nny = {p['ny']}\nreynolds = {p['reynolds']}\nlid_velocity = {p['lid']}\n"
                     f"cfl_limit = {p['cfl']}\nmax_steps = {p['max_steps']}\nmin_steps = {p['min_steps']}\n"
                     f"steady_tolerance = {p['steady_tolerance']}\npoisson_max_iterations = {p['maxit']}\n"
                     f"poisson_tolerance = {p['tol']}\nvalidate_reference = true\n")
        print("running solver_cavity to steady state", flush=True)
        r = subprocess.run([os.path.join(BIN, "solver_cavity"), cfg], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        if ": " in line:
            k, v = line.split(": ", 1)
            out[k] = v
    assert int(out["steps"]) == f.meta["step"]
    assert int(out["pressure_iterations"]) == f.meta["pressure_iterations"]
    assert out["benchmark_status"] == ("pass" if f.meta["validation_pass"] else "fail")
    assert abs(float(out["max_reference_relative_error"]) - f.meta["validation_max_relative_error"]) <= \
        1e-5 * max(f.meta["validation_max_relative_error"], 1e-12)
    for k in ["u_vertical_max", "u_vertical_min", "v_horizontal_max", "v_horizontal_min"]:
        assert abs(float(out[k]) - f.meta[k]) <= 1e-5 * abs(f.meta[k]), k
