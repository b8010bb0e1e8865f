"""GPU checks of the fused all-periodic multigrid level kernels against the unfused ones.

The fused down/up kernels evaluate the same per-cell expressions as the unfused level kernels
except that two intermediate means are not subtracted (the periodic operator annihilates
constants), so a solve must take the same number of iterations and agree to rounding. The
unfused path is selected by SOLVER_AMD_MG_UNFUSED=1 when a solver is built.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import solver_amd as sa
from tests.gpu_util import dev, host, rel_l2_active, zeros

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = sa.Context(0)
    yield c
    c.close()


def solve(ctx, g, rhs, precision, tol, unfused):
    os.environ["SOLVER_AMD_MG_UNFUSED"] = "1" if unfused else "0"
    try:
        o = sa.ProjectionOptions(1.0, 1.0, 400, tol)
        pb = sa.pressure_bc_set([sa.P_PERIODIC] * 6)
        sol = zeros(g, sa.CELL)
        d = sa.PoissonDiagnostics()
        ctx.run("of_solve_pressure_poisson", C.byref(g), C.byref(pb), C.byref(o), precision, sa.ptr(dev(rhs)),
                sa.ptr(sol), C.byref(d), None, 0)
    finally:
        os.environ.pop("SOLVER_AMD_MG_UNFUSED", None)
    return host(sol), d


# Equal spacings on every axis: with 2 pi / n spacings 96x48x32 is anisotropic enough that its
# residual stalls near the 1e-12 tolerance and rounding alone picks the iteration that first
# dips below it (66 fused, 67 unfused, 68 in the oracle).
@pytest.mark.parametrize("precision", [sa.FP64, sa.MIXED])
@pytest.mark.parametrize("shape", [(64, 64, 64), (96, 48, 32), (40, 24, 72)])
def test_fused_solve_matches_unfused(ctx, precision, shape):
    nx, ny, nz = shape
    g = sa.grid(nx, ny, nz, 2 * math.pi / 64, 2 * math.pi / 64, 2 * math.pi / 64)
    rng = np.random.default_rng(nx + ny + nz)
    rhs = rng.standard_normal(sa.storage_shape(nx, ny, nz, sa.CELL))
    tol = 1e-12 if precision == sa.FP64 else 1e-10
    a, da = solve(ctx, g, rhs, precision, tol, unfused=False)
    b, db = solve(ctx, g, rhs, precision, tol, unfused=True)
    assert da.converged == db.converged == 1
    assert abs(da.iterations - db.iterations) <= (0 if precision == sa.FP64 else 1), (da.iterations, db.iterations)
    # both solutions are within the tolerance of the exact one; they agree far better than that
    assert rel_l2_active(a, b, mean_removed=True) <= 10 * tol


def solve_env(ctx, g, rhs, precision, tol, env):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        o = sa.ProjectionOptions(1.0, 1.0, 400, tol)
        pb = sa.pressure_bc_set([sa.P_PERIODIC] * 6)
        sol = zeros(g, sa.CELL)
        d = sa.PoissonDiagnostics()
        ctx.run("of_solve_pressure_poisson", C.byref(g), C.byref(pb), C.byref(o), precision, sa.ptr(dev(rhs)),
                sa.ptr(sol), C.byref(d), None, 0)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return host(sol), d


# The marching level kernels come in 8- and 16-row tiles and the up kernel with two or three
# prefetch buffers; the one-shot small-level kernels keep their own 16-row tiles whatever the
# marching kernels use (their grid once followed the marching kernels' row count). Every variant
# takes the same iterations as the default and agrees with it to rounding.
@pytest.mark.parametrize("precision", [sa.FP64, sa.MIXED])
@pytest.mark.parametrize("shape", [(64, 64, 64), (128, 128, 128)])
@pytest.mark.parametrize("env", [{"SOLVER_AMD_FUSED_ROWS": "8,8,8,8"}, {"SOLVER_AMD_FUSED_ROWS": "16,8,16,8"},
                                 {"SOLVER_AMD_UP_BUFFERS": "3,2"}])
def test_fused_launch_variants_match_default(ctx, precision, shape, env):
    nx, ny, nz = shape
    g = sa.grid(nx, ny, nz, 2 * math.pi / nx, 2 * math.pi / ny, 2 * math.pi / nz)
    rng = np.random.default_rng(nx * 7 + ny + nz)
    rhs = rng.standard_normal(sa.storage_shape(nx, ny, nz, sa.CELL))
    tol = 1e-12 if precision == sa.FP64 else 1e-10
    a, da = solve_env(ctx, g, rhs, precision, tol, {})
    b, db = solve_env(ctx, g, rhs, precision, tol, env)
    assert da.converged == db.converged == 1
    assert abs(da.iterations - db.iterations) <= (0 if precision == sa.FP64 else 1), (da.iterations, db.iterations)
    assert rel_l2_active(a, b, mean_removed=True) <= 10 * tol


@pytest.mark.parametrize("precision", [sa.FP64, sa.MIXED])
def test_fused_stepper_matches_unfused(ctx, precision):
    n, steps = 48, 4
    L = 2 * math.pi
    dt, ns = sa.taylor_green_dt(n, n, n, 0.5, steps * 0.5 / (3 * n / L) * (1 - 1e-14))
    out = []
    for unfused in (False, True):
        os.environ["SOLVER_AMD_MG_UNFUSED"] = "1" if unfused else "0"
        try:
            st = sa.Stepper(ctx, sa.TAYLOR_GREEN, n, n, n, 1 / 1600, dt, 400, 1e-12, precision)
        finally:
            os.environ.pop("SOLVER_AMD_MG_UNFUSED", None)
        st.initialize(convention=1)
        its = [st.run(1).pressure_iterations for _ in range(ns)]
        vel, p, _, _, m = st.download()
        out.append((vel, p, its, m))
        st.close()
    (va, pa, ia, ma), (vb, pb, ib, mb) = out
    if precision == sa.FP64:
        assert ia == ib
    for x, y in zip(va, vb):
        assert rel_l2_active(x, y) <= 1e-11
    assert rel_l2_active(pa, pb, mean_removed=True) <= 1e-10
    assert ma.max_divergence_linf <= 1e-10


def solve_cg(ctx, g, pb, rhs, precision, tol, unfused):
    """Solve with the search direction and operator fused (default) or as two kernels."""
    os.environ["SOLVER_AMD_CG_UNFUSED"] = "1" if unfused else "0"
    try:
        o = sa.ProjectionOptions(1.0, 1.0, 600, tol)
        sol = zeros(g, sa.CELL)
        d = sa.PoissonDiagnostics()
        ctx.run("of_solve_pressure_poisson", C.byref(g), C.byref(pb), C.byref(o), precision, sa.ptr(dev(rhs)),
                sa.ptr(sol), C.byref(d), None, 0)
    finally:
        os.environ.pop("SOLVER_AMD_CG_UNFUSED", None)
    return host(sol), d


# periodic (lane-shuffle rows), ragged rows (nx not a multiple of 64), Neumann / Dirichlet faces
# (boundary images in x, y and z) and a single-plane grid
@pytest.mark.parametrize("precision", [sa.FP64, sa.MIXED])
@pytest.mark.parametrize("pbc,shape", [
    ([sa.P_PERIODIC] * 6, (64, 64, 64)),
    ([sa.P_PERIODIC] * 6, (96, 40, 24)),
    ([1, 1, 1, 1, 0, 0], (48, 40, 1)),
    ([2, 2, 1, 1, 0, 0], (32, 16, 24)),
    ([0, 1, 2, 2, 1, 0], (130, 20, 18)),
])
def test_fused_cg_matches_two_kernel_path(ctx, precision, pbc, shape):
    nx, ny, nz = shape
    h = (1.0 / nx, 1.3 / ny, 0.9 / nz if nz > 1 else 1.0)
    g = sa.grid(nx, ny, nz, *h)
    rng = np.random.default_rng(nx * ny + nz)
    rhs = rng.standard_normal(sa.storage_shape(nx, ny, nz, sa.CELL))
    pb = sa.pressure_bc_set(pbc, [0.2, -0.1, 0.0, 0.4, 0.1, 0.0], [0.3, 0.0, -0.2, 0.1, 0.0, 0.2])
    tol = 1e-12 if precision == sa.FP64 else 1e-10
    a, da = solve_cg(ctx, g, pb, rhs, precision, tol, unfused=False)
    b, db = solve_cg(ctx, g, pb, rhs, precision, tol, unfused=True)
    assert da.converged == db.converged == 1
    # d is formed by the same expression in both paths, so the iterations agree
    assert da.iterations == db.iterations, (da.iterations, db.iterations)
    assert rel_l2_active(a, b, mean_removed=all(t == sa.P_PERIODIC for t in pbc)) <= 1e-11


@pytest.mark.parametrize("precision", [sa.FP64, sa.MIXED])
@pytest.mark.parametrize("shape", [(64, 64, 64), (128, 64, 32), (128, 128, 128)])
def test_fused_solve_matches_oracle_and_fft(ctx, precision, shape):
    """The fused levels (64-wide tiles, several levels above the tail) against the oracle's
    restatement of the reference MGPCG and against the exact periodic solution (FFT): FP64
    iteration counts equal to the oracle's (within 3 on the anisotropic grid), and the solution
    within 1e-10 of the oracle's and 1e-9 of the exact one."""
    from oracle import ofref
    from tests.test_gpu_oracle import periodic_fft_solution

    nx, ny, nz = shape
    h = (2 * math.pi / nx, 2 * math.pi / ny, 2 * math.pi / nz)
    g = sa.grid(nx, ny, nz, *h)
    rng = np.random.default_rng(nx * 7 + nz)
    rhs = rng.standard_normal(sa.storage_shape(nx, ny, nz, sa.CELL))
    tol = 1e-12
    a, da = solve(ctx, g, rhs, precision, tol, unfused=False)
    og = ofref.make_grid(nx, ny, nz, *h)
    ref, rd, _ = ofref.poisson(og, ofref.pbcset([ofref.P_PERIODIC] * 6), ofref.opts(1.0, 1.0, 400, tol), rhs.copy())
    assert da.converged == rd.converged == 1
    if precision == sa.FP64:
        # equal counts on the isotropic grid; on the anisotropic one (spacing 1:2:4) the CG
        # amplifies the rounding of differently ordered sums (residual histories agree to 1e-15
        # at first, then drift apart about tenfold per iteration), so the stop can move by a few
        # iterations while the solutions still agree to 1e-10
        assert abs(da.iterations - rd.iterations) <= (0 if nx == ny == nz else 3), (da.iterations, rd.iterations)
    act = a[1:-1, 1:-1, 1:-1]
    exact = periodic_fft_solution(rhs[1:-1, 1:-1, 1:-1], h)
    scale = np.max(np.abs(exact))
    assert np.max(np.abs(act - ref[1:-1, 1:-1, 1:-1])) <= 1e-10 * scale
    assert np.max(np.abs(act - exact)) <= 1e-9 * scale


def solve_ring(ctx, g, pb, rhs, p0, precision, ring):
    """A solve with `ring` search-direction slots ("" for the default)."""
    os.environ["SOLVER_AMD_CG_RING"] = ring
    try:
        o = sa.ProjectionOptions(1.0, 1.0, 600, 1e-12)
        sol = dev(p0)
        d = sa.PoissonDiagnostics()
        ctx.run("of_solve_pressure_poisson", C.byref(g), C.byref(pb), C.byref(o), precision, sa.ptr(dev(rhs)),
                sa.ptr(sol), C.byref(d), None, 0)
    finally:
        os.environ.pop("SOLVER_AMD_CG_RING", None)
    return host(sol), d.iterations


# The pressure is accumulated from the kept search directions in one pass at the end; with a
# ring shorter than the iteration count the oldest directions are folded into p every R
# iterations. Each cell sees the same sequence of additions either way, so the results are
# bitwise equal, from a zero and from a nonzero warm start.
@pytest.mark.parametrize("precision", [sa.FP64, sa.MIXED])
@pytest.mark.parametrize("pbc,shape", [([sa.P_PERIODIC] * 6, (64, 48, 40)), ([1, 1, 0, 0, 2, 2], (72, 40, 32))])
def test_direction_ring_folds_are_bitwise_neutral(ctx, precision, pbc, shape):
    nx, ny, nz = shape
    g = sa.grid(nx, ny, nz, 1.0 / nx, 1.1 / ny, 0.9 / nz)
    pb = sa.pressure_bc_set(pbc)
    rng = np.random.default_rng(11)
    shp = sa.storage_shape(nx, ny, nz, sa.CELL)
    rhs = rng.standard_normal(shp)
    for p0 in (np.zeros(shp), rng.standard_normal(shp)):
        ref, its = solve_ring(ctx, g, pb, rhs, p0, precision, "")
        assert its > 6  # several folds with 2 and 3 slots
        for ring in ("2", "3"):
            sol, its_r = solve_ring(ctx, g, pb, rhs, p0, precision, ring)
            assert its_r == its
            assert np.array_equal(sol, ref)


def solve_lds(ctx, g, rhs, p0, precision, lds):
    """A periodic solve with the LDS-staged CG passes on (lds=True) or the marching ones."""
    os.environ["SOLVER_AMD_CG_LDS"] = "1" if lds else "0"
    try:
        o = sa.ProjectionOptions(1.0, 1.0, 400, 1e-12 if precision == sa.FP64 else 1e-10)
        pb = sa.pressure_bc_set([sa.P_PERIODIC] * 6)
        sol = dev(p0)
        d = sa.PoissonDiagnostics()
        ctx.run("of_solve_pressure_poisson", C.byref(g), C.byref(pb), C.byref(o), precision, sa.ptr(dev(rhs)),
                sa.ptr(sol), C.byref(d), None, 0)
    finally:
        os.environ.pop("SOLVER_AMD_CG_LDS", None)
    return host(sol), d


# On periodic grids whose 64 x 4 tiles and 8-plane ranges are all full (and at least 1024 of
# them), the search-direction and residual passes stage the lateral neighbours in LDS. They form
# every value by the marching kernels' expressions with the same tiles and per-thread order, so
# the solves are bitwise those of the marching path (SOLVER_AMD_CG_LDS=0), iteration for
# iteration, from a zero and from a nonzero warm start.
@pytest.mark.parametrize("precision", [sa.FP64, sa.MIXED])
@pytest.mark.parametrize("shape,h", [((128, 128, 128), (1.0, 1.0, 1.0)), ((64, 256, 128), (1.0, 0.7, 1.3))])
def test_lds_cg_passes_are_bitwise_the_marching_ones(ctx, precision, shape, h):
    nx, ny, nz = shape
    assert (nx // 64) * (ny // 4) * (nz // 8) >= 1024
    g = sa.grid(nx, ny, nz, h[0] / nx, h[1] / ny, h[2] / nz)
    rng = np.random.default_rng(nx + 3 * ny + 7 * nz)
    shp = sa.storage_shape(nx, ny, nz, sa.CELL)
    rhs = rng.standard_normal(shp)
    for p0 in (np.zeros(shp), rng.standard_normal(shp)):
        a, da = solve_lds(ctx, g, rhs, p0, precision, True)
        b, db = solve_lds(ctx, g, rhs, p0, precision, False)
        assert da.converged == db.converged == 1
        assert da.iterations == db.iterations > 5
        assert da.relative_residual == db.relative_residual
        assert np.array_equal(a, b)
