"""Pins the CPU restatement (oracle/ofref.cpp) against the reference's own outputs.

Every fixture under tests/golden/fixtures/ was produced by the reference CPU sources
(tests/golden/make_golden.py). The restatement must reproduce them to rounding level;
iteration counts of the MGPCG solves must match exactly.
"""
import numpy as np
import pytest

from oracle import ofref
from tests.golden import fixtures as fx

TIGHT = 1e-13  # relative max-abs error allowed for the restatement (rounding level)


def close(a, b, tol=TIGHT, what=""):
    scale = max(np.max(np.abs(b)), 1e-300)
    err = np.max(np.abs(a - b)) / scale
    assert err <= tol, f"{what}: rel max err {err:.3e} > {tol:.1e}"


def grid_of(p):
    return ofref.make_grid(p["nx"], p["ny"], p["nz"], p["dx"], p["dy"], p["dz"])


def bcs_of(p):
    t = p["bc"]
    vel = np.array(p.get("bcvel", [0.0] * 18)).reshape(6, 3)
    pr = p.get("bcp", [0.0] * 6)
    return ofref.bcset(t, vel, pr)


def pbcs_of(p):
    return ofref.pbcset(p["pbc"], p.get("pbcval"), p.get("pbcgrad"))


def vel_in(f, pre=""):
    return [f.inp[pre + "u"].copy(), f.inp[pre + "v"].copy(), f.inp[pre + "w"].copy()]


@pytest.mark.parametrize("name", fx.names("operators_"))
def test_operators(name):
    f = fx.load(name)
    g = grid_of(f.params)
    vel = vel_in(f)
    gr = ofref.gradient(g, f.inp["p"].copy())
    for a, k in zip(gr, ["gx", "gy", "gz"]):
        close(a, f.out[k], what=k)
    close(ofref.divergence(g, vel), f.out["div"], what="div")
    close(ofref.laplacian(g, ofref.CELL, f.inp["p"].copy()), f.out["lap_p"], what="lap_p")
    for loc, k in zip([1, 2, 3], ["u", "v", "w"]):
        close(ofref.laplacian(g, loc, f.inp[k].copy()), f.out["lap_" + k], what="lap_" + k)


@pytest.mark.parametrize("name", fx.names("advection_"))
def test_advection(name):
    f = fx.load(name)
    out = ofref.advection(grid_of(f.params), vel_in(f), f.params["scheme"])
    for a, k in zip(out, ["au", "av", "aw"]):
        close(a, f.out[k], what=k)


def test_diffusion_and_cfl():
    f = fx.load("diffusion_3d")
    out = ofref.diffusion(grid_of(f.params), f.params["nu"], vel_in(f))
    for a, k in zip(out, ["du", "dv", "dw"]):
        close(a, f.out[k], what=k)
    f = fx.load("cfl_3d")
    c = ofref.cfl(grid_of(f.params), vel_in(f), f.params["dt"])
    for k in ["dt", "max_cfl", "max_u", "max_v", "max_w"]:
        assert abs(c[k] - f.meta[k]) <= TIGHT * abs(f.meta[k]), k


@pytest.mark.parametrize("name", fx.names("vbc_"))
def test_velocity_bc(name):
    f = fx.load(name)
    vel = vel_in(f)
    ofref.velocity_bc(grid_of(f.params), bcs_of(f.params), vel)
    for a, k in zip(vel, "uvw"):
        np.testing.assert_array_equal(a, f.out[k])


@pytest.mark.parametrize("name", fx.names("pbc_"))
def test_pressure_bc(name):
    f = fx.load(name)
    p = f.inp["p"].copy()
    ofref.pressure_bc(grid_of(f.params), pbcs_of(f.params), p)
    np.testing.assert_array_equal(p, f.out["p"])


@pytest.mark.parametrize("name", fx.names("tpbc_"))
def test_total_pressure_bc(name):
    f = fx.load(name)
    p = f.inp["p"].copy()
    ofref.total_pressure_bc(grid_of(f.params), bcs_of(f.params), vel_in(f, "s"), p)
    np.testing.assert_array_equal(p, f.out["p"])


@pytest.mark.parametrize("name", fx.names("cavity_rhs_"))
def test_cavity_predictor_rhs(name):
    f = fx.load(name)
    p = f.params
    prev = vel_in(f, "p") if p["has_prev"] else None
    adv, rhs = ofref.predictor_rhs(grid_of(p), vel_in(f), f.inp["p"].copy(), prev, p["nu"], p["dt"], cavity=True)
    for a, k in zip(adv + rhs, ["au", "av", "aw", "ru", "rv", "rw"]):
        close(a, f.out[k], what=k)


@pytest.mark.parametrize("name", fx.names("adi_"))
def test_predictor_adi(name):
    f = fx.load(name)
    out, ls = ofref.predictor_adi(grid_of(f.params), bcs_of(f.params), f.params["alpha"], vel_in(f))
    assert ls == f.meta["line_solves"]
    for a, k in zip(out, "uvw"):
        close(a, f.out[k], what=k)


@pytest.mark.parametrize("name", fx.names("bprhs_"))
def test_build_pressure_rhs(name):
    f = fx.load(name)
    p = f.params
    o = ofref.opts(p["dt"], p["density"])
    close(ofref.build_pressure_rhs(grid_of(p), bcs_of(p), o, vel_in(f)), f.out["rhs"], what="rhs")


@pytest.mark.parametrize("name", fx.names("correct_"))
def test_correct_velocity(name):
    f = fx.load(name)
    p = f.params
    o = ofref.opts(p["dt"], p["density"])
    out = ofref.correct_velocity(grid_of(p), bcs_of(p), o, vel_in(f), f.inp["phi"].copy())
    for a, k in zip(out, "uvw"):
        close(a, f.out[k], what=k)


@pytest.mark.parametrize("name", fx.names("project_"))
def test_project_velocity(name):
    f = fx.load(name)
    p = f.params
    o = ofref.opts(p["dt"], p["density"], p["maxit"], p["tol"])
    phi, corr, prhs, d = ofref.project_velocity(grid_of(p), bcs_of(p), o, vel_in(f), f.inp["phi"].copy())
    assert d.poisson.iterations == f.meta["iterations"]
    assert d.poisson.converged == f.meta["converged"]
    close(phi, f.out["phi"], 1e-12, "phi")
    for a, k in zip(corr, "uvw"):
        close(a, f.out[k], 1e-12, k)
    close(prhs, f.out["prhs"], what="prhs")
    for k in ["rhs_l2", "divergence_l2_before"]:
        assert abs(getattr(d, k) - f.meta[k]) <= 1e-12 * abs(f.meta[k]), k
    assert abs(d.divergence_l2_after - f.meta["divergence_l2_after"]) <= 1e-12 * max(1.0, f.meta["rhs_l2"])


@pytest.mark.parametrize("name", [n for n in fx.names("poisson_") if not n.startswith("poisson_rhs")])
def test_poisson(name):
    f = fx.load(name)
    p = f.params
    o = ofref.opts(1.0, 1.0, p["maxit"], p["tol"])
    sol, d, hist = ofref.poisson(grid_of(p), pbcs_of(p), o, f.inp["rhs"].copy(), f.inp["p"].copy())
    assert d.iterations == f.meta["iterations"]
    assert d.multigrid_levels == f.meta["multigrid_levels"]
    assert d.coarse_unknowns == f.meta["coarse_unknowns"]
    assert d.zero_mean_enforced == f.meta["zero_mean_enforced"]
    assert d.converged == f.meta["converged"] == 1
    close(sol, f.out["p"], 1e-12, "p")
    ref_hist = f.arrays["residual_history"]
    assert hist.shape == ref_hist.shape
    np.testing.assert_allclose(hist, ref_hist, rtol=1e-9, atol=1e-9 * ref_hist[0])


@pytest.mark.parametrize("name", fx.names("poisson_rhs_"))
def test_poisson_rhs_from_pressure(name):
    f = fx.load(name)
    p = f.params
    close(ofref.poisson_rhs_from_pressure(grid_of(p), pbcs_of(p), f.inp["p"].copy()), f.out["rhs"], what="rhs")


def tg_cfg(p):
    return ofref.tg_config(p["nx"], p["ny"], p["nz"], p["viscosity"], p["cfl"], p["final_time"], p["maxit"], p["tol"])


@pytest.mark.parametrize("name", [n for n in fx.names("tg") if n.endswith("_init")])
def test_tg_init(name):
    f = fx.load(name)
    c = tg_cfg(f.params)
    s = ofref.tg_init(c)
    assert abs(s.metrics.dt - f.meta["dt"]) <= 1e-15 * f.meta["dt"]
    for a, k in zip(s.vel + [s.p], "uvwp"):
        close(a, f.out[k], 1e-14, k)


@pytest.mark.parametrize("name", [n for n in fx.names("tg") if n.endswith("_steps")])
def test_tg_steps(name):
    f = fx.load(name)
    p = f.params
    c = tg_cfg(p)
    if f.inp:
        g = ofref.tg_grid(c)
        s = ofref.State(g)
        s.vel = vel_in(f)
        s.p = f.inp["p"].copy()
        s.adv = vel_in(f, "a")
        s.has_prev.value = p["has_prev"]
        s.metrics.dt = ofref.tg_dt(c)
    else:
        s = ofref.tg_init(c)
    ofref.tg_steps(c, s, p["steps"])
    assert s.metrics.step == f.meta["step"]
    assert s.metrics.pressure_iterations == f.meta["pressure_iterations"]
    assert abs(s.metrics.time - f.meta["time"]) <= 1e-14
    for a, k in zip(s.vel + [s.p] + s.adv, ["u", "v", "w", "p", "au", "av", "aw"]):
        close(a, f.out[k], 1e-12, k)


def test_tg_nonconverged_message():
    f = fx.load("tg_nonconverged")
    c = tg_cfg(f.params)
    s = ofref.tg_init(c)
    with pytest.raises(ofref.OracleError) as e:
        ofref.tg_steps(c, s, 1)
    assert "Taylor-Green pressure solve did not converge" in f.meta["error"]
    assert str(e.value) == f.meta["error"]


def test_cavity_init_and_steps():
    f = fx.load("cavity16_init")
    p = f.params
    c = ofref.cavity_config(p["nx"], p["ny"], p["reynolds"], p["lid"], p["cfl"], max_iterations=p["maxit"],
                            tolerance=p["tol"])
    s = ofref.cavity_init(c)
    for a, k in zip(s.vel + [s.p], "uvwp"):
        np.testing.assert_array_equal(a, f.out[k])
    f = fx.load("cavity16_steps")
    ofref.cavity_steps(c, s, f.params["steps"])
    assert s.metrics.step == f.meta["step"]
    assert s.metrics.pressure_iterations == f.meta["pressure_iterations"]
    for a, k in zip(s.vel + [s.p] + s.adv, ["u", "v", "w", "p", "au", "av", "aw"]):
        close(a, f.out[k], 1e-12, k)


@pytest.mark.parametrize("name", ["channel_couette_32", "channel_poiseuille_32", "channel_couette_128",
                                  "channel_poiseuille_128"])
def test_channel_flow(name):
    """run_channel_flow (solver/channel_flow.cpp:432-541): the reference's final metrics, its
    x-averaged streamwise profile and its validation figure, from the reference itself."""
    f = fx.load(name)
    p = f.params
    c = ofref.channel_config(p["case"], p["nx"], p["ny"], p["viscosity"], p["top"], p["drop"], p["cfl"], p["steps"],
                             p["maxit"], p["tol"])
    s = ofref.channel_init(c)
    ofref.channel_steps(c, s, p["steps"])
    m = s.metrics
    assert m.step == f.meta["step"] == p["steps"]
    assert m.pressure_iterations == f.meta["pressure_iterations"]
    assert m.dt == f.meta["dt"]
    for k in ["max_cfl", "max_velocity_change"]:
        assert abs(getattr(m, k) - f.meta[k]) <= 1e-12 * abs(f.meta[k]), k
    assert abs(m.divergence_l2 - f.meta["divergence_l2"]) <= 1e-3 * f.meta["divergence_l2"] + 1e-16
    prof, err = ofref.channel_profile(c, s.vel[0])
    close(prof, f.out["profile"], 1e-13, "profile")
    assert abs(err - f.meta["relative_l2_error"]) <= 1e-9 * f.meta["relative_l2_error"]
    # the reference's own acceptance (solver/channel_flow.cpp:421-430)
    assert f.meta["pass"] == 1 and err <= 5e-3 and m.divergence_l2 <= 1e-10


def test_channel_flow_rejects_nonconverged_projection():
    f = fx.load("channel_nonconverged")
    p = f.params
    c = ofref.channel_config(p["case"], p["nx"], p["ny"], p["viscosity"], p["top"], p["drop"], p["cfl"], p["steps"],
                             p["maxit"], p["tol"])
    s = ofref.channel_init(c)
    with pytest.raises(RuntimeError) as e:
        ofref.channel_steps(c, s, p["steps"])
    assert str(e.value) == f.meta["error"]


@pytest.mark.parametrize("n", [16, 32, 64])
def test_mms(n):
    f = fx.load(f"mms_{n}")
    r = ofref.mms(n)
    for k in ["gradient_error", "divergence_error", "laplacian_error"]:
        assert abs(r[k] - f.meta[k]) <= 1e-12 * f.meta[k], k
