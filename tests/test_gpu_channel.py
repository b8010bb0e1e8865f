"""GPU parity of the channel-flow cases (reference solver/channel_flow.cpp) against the oracle's
restatement and the reference's own run_channel_flow results (golden fixtures).

* Couette: periodic x, no-slip lower wall, upper wall moving at top_velocity.
* Poiseuille: fixed pressure drop along x (Dirichlet pressure correction in the multigrid),
  no-slip walls.

Both start from the analytic steady profile; the device stepper must reproduce the oracle's
fields at 1e-10 with equal pressure iterations, the reference's x-averaged profile, metrics and
validation figure, and its non-convergence message.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import solver_amd as sa
from oracle import ofref
from tests.golden import fixtures as fx
from tests.gpu_util import dev, host, rel_max, zeros

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "solver_amd", "cpp", "build", "bin")
CFG = os.path.join(ROOT, "solver_amd", "cpp", "configs")
IT = 1e-10  # iterative-solver parity bar (MGPCG tolerance 1e-10)
CASES = {"couette": sa.CHANNEL_COUETTE, "poiseuille": sa.CHANNEL_POISEUILLE}


@pytest.fixture(scope="module")
def ctx():
    c = sa.Context(0)
    yield c
    c.close()


def oracle_config(p, **over):
    q = dict(p, **over)
    return ofref.channel_config(q["case"], q["nx"], q["ny"], q["viscosity"], q["top"], q["drop"], q["cfl"],
                                q["steps"], q["maxit"], q["tol"])


def device_stepper(ctx, p, **over):
    q = dict(p, **over)
    kind = CASES[q["case"]]
    dt = sa.channel_flow_dt(kind, q["nx"], q["viscosity"], q["top"], q["drop"], q["cfl"])
    print(f"channel {q['case']} {q['nx']}x{q['ny']} steps={q['steps']} dt={dt}", flush=True)
    return sa.Stepper(ctx, kind, q["nx"], q["ny"], 1, q["viscosity"], dt, q["maxit"], q["tol"], sa.FP64, 0,
                      q["top"], q["drop"]), dt


def x_averaged_profile(u):
    """Sum of the active x faces of every active row of plane k = 1, divided by their count (the
    reference's summation order, solver/channel_flow.cpp:258-279)."""
    a = np.asarray(u)[1, 1:-1, 1:-1]
    out = np.zeros(a.shape[0])
    for i in range(a.shape[1]):
        out += a[:, i]
    return out / a.shape[1]


@pytest.mark.parametrize("case", ["couette", "poiseuille"])
def test_channel_init_matches_oracle(ctx, case):
    p = fx.load(f"channel_{case}_32").params
    st, dt = device_stepper(ctx, p)
    c = oracle_config(p)
    ref = ofref.channel_init(c)
    assert dt == ref.metrics.dt
    st.initialize(0)
    vel, pt, _, _, _ = st.download()
    for a, b, k in zip(vel + [pt], ref.vel + [ref.p], "uvwp"):
        err = rel_max(a, b) if np.any(b) else float(np.max(np.abs(a)))
        assert err <= 1e-14, f"{k}: {err:.3e}"
    st.close()


@pytest.mark.parametrize("name", ["channel_couette_32", "channel_poiseuille_32", "channel_couette_128",
                                  "channel_poiseuille_128"])
def test_channel_steps_vs_oracle_and_reference(ctx, name):
    f = fx.load(name)
    p = f.params
    st, dt = device_stepper(ctx, p)
    st.initialize(0)
    m = st.run(p["steps"])
    vel, pt, adv, _, _ = st.download()
    st.close()
    c = oracle_config(p)
    ref = ofref.channel_init(c)
    ofref.channel_steps(c, ref, p["steps"])
    # oracle: whole fields and the step metrics
    assert m.step == ref.metrics.step == p["steps"]
    assert m.pressure_iterations == ref.metrics.pressure_iterations == f.meta["pressure_iterations"]
    for a, b, k in zip(vel + [pt] + adv, ref.vel + [ref.p] + ref.adv, ["u", "v", "w", "p", "au", "av", "aw"]):
        err = rel_max(a, b) if np.any(b) else float(np.max(np.abs(a)))
        assert err <= IT, f"{k}: {err:.3e}"
    for k in ["time", "dt", "max_cfl", "max_velocity_change"]:
        assert abs(getattr(m, k) - f.meta[k]) <= 1e-9 * abs(f.meta[k]), k
    assert m.divergence_l2 <= 1e-10 and f.meta["divergence_l2"] <= 1e-10
    # the reference's own profile and validation figure
    prof = x_averaged_profile(vel[0])
    assert rel_max(prof, f.out["profile"]) <= IT
    _, err = ofref.channel_profile(c, vel[0])
    assert abs(err - f.meta["relative_l2_error"]) <= 1e-6 * f.meta["relative_l2_error"]
    assert err <= 5e-3


def test_channel_rejects_nonconverged_projection(ctx):
    f = fx.load("channel_nonconverged")
    p = f.params
    st, _ = device_stepper(ctx, p)
    st.initialize(0)
    with pytest.raises(sa.SolverError) as e:
        st.run(p["steps"])
    st.close()
    want = f.meta["error"]
    assert "channel flow pressure solve did not converge at step 1 (iterations=1" in str(e.value)
    assert str(e.value).split(", relative_residual")[0] == want.split(", relative_residual")[0]


def test_periodic_adi_preserves_couette_state(ctx):
    """Linear u = U y meets both wall relations exactly and has zero discrete Laplacian, so the
    Helmholtz solve returns it unchanged (tests/test_runtime.cpp:877-915)."""
    nx, ny, top = 16, 12, 1.0
    g = sa.grid(nx, ny, 1, 1.0 / nx, 1.0 / ny, 1.0)
    t = [sa.PERIODIC, sa.PERIODIC, sa.NO_SLIP, sa.PRESCRIBED, sa.SYMMETRY, sa.SYMMETRY]
    vel = np.zeros((6, 3))
    vel[3] = [top, 0.0, 0.0]
    b = sa.bc_set(t, vel)
    ob = ofref.bcset(t, vel)
    og = ofref.make_grid(nx, ny, 1, 1.0 / nx, 1.0 / ny, 1.0)
    shp = [sa.storage_shape(nx, ny, 1, loc) for loc in (sa.FACE_X, sa.FACE_Y, sa.FACE_Z)]
    u = np.zeros(shp[0])
    y = (np.arange(shp[0][1]) - 1 + 0.5) / ny
    u[:] = top * y[None, :, None]
    rhs = [u, np.zeros(shp[1]), np.zeros(shp[2])]
    ofref.velocity_bc(og, ob, rhs)
    out = [zeros(g, loc) for loc in (sa.FACE_X, sa.FACE_Y, sa.FACE_Z)]
    ls = C.c_int()
    ctx.run("of_solve_predictor_adi", C.byref(g), C.byref(b), C.c_double(0.01), *[sa.ptr(dev(a)) for a in rhs],
            *map(sa.ptr, out), C.byref(ls))
    assert ls.value > 0
    pu = host(out[0])[1:-1, 1:-1, 1:-1]
    ru = rhs[0][1:-1, 1:-1, 1:-1]
    assert np.sqrt(np.sum((pu - ru) ** 2) / np.sum(ru ** 2)) <= 1e-12
    assert max(float(np.max(np.abs(host(o)[1:-1, 1:-1, 1:-1]))) for o in out) <= top + 1e-12


def run_tool(*args, timeout=300):
    print(f"running solver_channel {' '.join(map(str, args))}", flush=True)
    return subprocess.run([os.path.join(BIN, "solver_channel"), *map(str, args)], capture_output=True, text=True,
                          timeout=timeout)


def kv(text):
    out = {}
    for line in text.splitlines():
        if ": " in line:
            k, v = line.split(": ", 1)
            try:
                out[k] = float(v)
            except ValueError:
                out[k] = v
    return out


@pytest.mark.parametrize("case,size", [("couette", "smoke"), ("poiseuille", "smoke"), ("couette", "128"),
                                       ("poiseuille", "128")])
def test_channel_tool_vs_reference(case, size):
    n = 32 if size == "smoke" else 128
    f = fx.load(f"channel_{case}_{n}")
    r = run_tool(os.path.join(CFG, f"channel_{case}_{size}.cfg"))
    assert r.returncode == 0, r.stdout + r.stderr
    out = kv(r.stdout)
    assert out["config"].startswith(f"case={case}, nx={n}, ny={n}, viscosity=0.1")
    assert int(out["steps"]) == f.meta["step"]
    assert int(out["pressure_iterations"]) == f.meta["pressure_iterations"]
    for k in ["time", "dt", "max_cfl", "max_velocity_change"]:
        assert abs(out[k] - f.meta[k]) <= 1e-5 * abs(f.meta[k]), k  # printed with 6 digits
    assert out["reference_dataset"] == f.meta["reference_dataset"]
    assert abs(out["profile_relative_l2_error"] - f.meta["relative_l2_error"]) <= 1e-5 * f.meta["relative_l2_error"]
    assert out["benchmark_status"] == "pass"


def test_channel_tool_reports_nonconvergence(tmp_path):
    cfg = tmp_path / "bad.cfg"
    cfg.write_text("case = poiseuille\nnx = 32\nny = 32\nsteps = 4\npoisson_max_iterations = 1\n"
                   "poisson_tolerance = 1e-30\n")
    r = run_tool(cfg)
    assert r.returncode == 1
    assert "solver_channel failed: channel flow pressure solve did not converge at step 1" in r.stderr
