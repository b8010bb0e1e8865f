"""Known-answer checks of the device path, written anew after the reference's own tests
(tests/test_runtime.cpp), at the reference's sizes and at production sizes the golden fixtures
never reach. Each answer is fixed by mathematics, not by the oracle's restatement:

* ADI predictor = the inverse of the factorised operator (1 - a dxx)(1 - a dyy)(1 - a dzz),
  built as a dense matrix from basis vectors with the reference's ghost relations
  (tests/test_runtime.cpp:1009-1056), and on periodic grids the FFT diagonalisation
  prod_a (1 + a (2 - 2 cos theta_a) / h_a^2);
* MGPCG recovers the exact discrete solution of Dirichlet, pure-Neumann and mixed problems,
  obtained from sine / cosine transforms (the eigenvectors of the cell-centred 3-point operator
  with ghost = -interior or ghost = interior), at 32^2 (the reference's case, :1178-1269),
  256^2 and 64^3;
* TVD advection of a step keeps the explicit update within the data's bounds (:1414-1445);
* a quiescent projection takes 0 iterations and returns zeros (:1447-1477);
* projecting a pure gradient field recovers the pressure and leaves zero velocity (:1479-1539).
"""
import ctypes as C
import math

import numpy as np
import pytest
import scipy.fft as sfft

import solver_amd as sa
from oracle import ofref
from tests.gpu_util import dev, host, zeros, zeros_vel

pytestmark = pytest.mark.gpu

NS, PV, SYM, PER = sa.NO_SLIP, sa.PRESCRIBED, sa.SYMMETRY, sa.PERIODIC


@pytest.fixture(scope="module")
def ctx():
    c = sa.Context(0)
    yield c
    c.close()


def coords(g, loc):
    """Storage-point coordinates (z, y, x mesh) of a field (core/field_layout.hpp:227-238)."""
    n = [g.nx, g.ny, g.nz]
    h = [g.dx, g.dy, g.dz]
    out = []
    for a in range(3):
        cnt = n[a] + 2 + (1 if loc == a + 1 else 0)
        idx = np.arange(cnt) - 1.0
        out.append(idx * h[a] if loc == a + 1 else (idx + 0.5) * h[a])
    z, y, x = np.meshgrid(out[2], out[1], out[0], indexing="ij")
    return x, y, z


# ---------------------------------------------------------------- ADI vs dense factorised operator
def ofref_grid(g):
    return ofref.make_grid(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz)


def helmholtz_axis(f, alpha, axis, h):
    """out = f - alpha * d2f/dx_axis^2 on the active range; elsewhere a copy of f."""
    out = f.copy()
    ax = 2 - axis  # numpy axis of (z, y, x) storage
    sl = [slice(1, -1)] * 3
    lo, hi = list(sl), list(sl)
    lo[ax] = slice(0, -2)
    hi[ax] = slice(2, None)
    out[tuple(sl)] = f[tuple(sl)] - alpha * (f[tuple(hi)] - 2 * f[tuple(sl)] + f[tuple(lo)]) / (h * h)
    return out


def factorised_apply(g, bc_types, comp, field, alpha):
    """The reference test's apply_factorized_operator: BC fill, then one axis at a time."""
    og = ofref_grid(g)
    ob = ofref.bcset(bc_types)
    vel = ofref.zeros_vel(og)
    vel[comp] = field.copy()
    h = [g.dx, g.dy, g.dz]
    for axis in range(3):
        ofref.velocity_bc(og, ob, vel)
        vel[comp] = helmholtz_axis(vel[comp], alpha, axis, h[axis])
    ofref.velocity_bc(og, ob, vel)
    return vel[comp]


def unknowns(g, bc_types, comp):
    """Active entries of one component minus faces fixed by a wall (no-slip, prescribed,
    symmetry) on its normal axis; on a periodic normal axis the duplicated last face is the
    first one's copy, not an unknown (tests/test_runtime.cpp:501-538)."""
    shp = sa.storage_shape(g.nx, g.ny, g.nz, comp + 1)
    rng = [range(1, shp[2] - 1), range(1, shp[1] - 1), range(1, shp[0] - 1)]
    lo_f, hi_f = bc_types[2 * comp], bc_types[2 * comp + 1]
    r = list(rng[comp])
    if lo_f in (NS, PV, SYM):
        r = r[1:]
    if hi_f in (NS, PV, SYM) or hi_f == PER:
        r = r[:-1]
    rng[comp] = r
    return [(k, j, i) for k in rng[2] for j in rng[1] for i in rng[0]], shp


def basis(shp, comp, idx, g, bc_types):
    e = np.zeros(shp)
    e[idx] = 1.0
    if bc_types[2 * comp] == PER:  # the periodic duplicate carries the same value
        n = [g.nx, g.ny, g.nz][comp]
        k, j, i = idx
        dup = [i, j, k]
        if dup[comp] == 1:
            dup[comp] = n + 1
            e[dup[2], dup[1], dup[0]] = 1.0
    return e


ADI_CASES = [
    # the reference's case: 4 x 3 x 1, no-slip walls, symmetry in z, alpha 0.0375
    ((4, 3, 1), (0.25, 1 / 3, 1.0), [NS, NS, NS, NS, SYM, SYM], 0.0375),
    ((6, 5, 4), (0.2, 0.25, 0.3), [NS, SYM, PV, NS, SYM, NS], 0.02),
    ((8, 6, 1), (0.125, 1 / 6, 1.0), [PER, PER, NS, PV, SYM, SYM], 0.05),
    ((5, 6, 4), (0.2, 0.15, 0.25), [SYM, NS, PER, PER, NS, PV], 0.01),
]


@pytest.mark.parametrize("shape,h,bc_types,alpha", ADI_CASES)
def test_adi_matches_dense_factorised_operator(ctx, shape, h, bc_types, alpha):
    g = sa.grid(*shape, *h)
    b = sa.bc_set(bc_types)
    rng = np.random.default_rng(3)
    rhs = [rng.standard_normal(sa.storage_shape(*shape, loc)) for loc in (sa.FACE_X, sa.FACE_Y, sa.FACE_Z)]
    # the reference's right-hand side for u on its own case
    if shape == (4, 3, 1):
        x, y, _ = coords(g, sa.FACE_X)
        rhs[0] = 0.2 + 0.4 * x - 0.3 * y + 0.1 * x * y
    og, ob = ofref_grid(g), ofref.bcset(bc_types)
    ofref.velocity_bc(og, ob, rhs)
    out = zeros_vel(g)
    ls = C.c_int()
    ctx.run("of_solve_predictor_adi", C.byref(g), C.byref(b), C.c_double(alpha), *map(sa.ptr, map(dev, rhs)),
            *map(sa.ptr, out), C.byref(ls))
    assert ls.value > 0
    for comp in range(3):
        if shape[2] == 1 and comp == 2:
            continue
        unk, shp = unknowns(g, bc_types, comp)
        m = np.zeros((len(unk), len(unk)))
        for col, idx in enumerate(unk):
            img = factorised_apply(g, bc_types, comp, basis(shp, comp, idx, g, bc_types), alpha)
            m[:, col] = [img[u] for u in unk]
        want = np.linalg.solve(m, np.array([rhs[comp][u] for u in unk]))
        got = np.array([host(out[comp])[u] for u in unk])
        assert np.max(np.abs(got - want)) <= 1e-10 * max(1.0, np.max(np.abs(want))), comp


def periodic_helmholtz_fft(r, h, alpha):
    """Solution of prod_a (1 - alpha d_aa) x = r on the torus (r of shape (nz, ny, nx))."""
    nz, ny, nx = r.shape
    th = [2 * np.pi * np.fft.fftfreq(n) for n in (nx, ny, nz)]
    lam = [1 + alpha * (2 - 2 * np.cos(t)) / (hh * hh) for t, hh in zip(th, h)]
    den = lam[2][:, None, None] * lam[1][None, :, None] * lam[0][None, None, :]
    return np.real(np.fft.ifftn(np.fft.fftn(r) / den))


def torus(a, comp):
    """Active torus values of a face component (drops the duplicated last face)."""
    x = a[1:-1, 1:-1, 1:-1]
    if comp == 0:
        return x[:, :, :-1]
    if comp == 1:
        return x[:, :-1, :]
    return x[:-1, :, :]


def periodic_field(rng, shape, comp):
    nx, ny, nz = shape
    t = rng.standard_normal((nz, ny, nx))
    full = np.zeros(sa.storage_shape(nx, ny, nz, comp + 1))
    a = full[1:-1, 1:-1, 1:-1]
    if comp == 0:
        a[:, :, :-1] = t
        a[:, :, -1] = t[:, :, 0]
    elif comp == 1:
        a[:, :-1, :] = t
        a[:, -1, :] = t[:, 0, :]
    else:
        a[:-1, :, :] = t
        a[-1, :, :] = t[0, :, :]
    return full, t


@pytest.mark.parametrize("solver", ["adi", "cg"])
@pytest.mark.parametrize("shape,alpha", [
    ((16, 12, 8), 0.02),        # line solves (taps decay too slowly for a short convolution)
    ((256, 64, 32), 3e-4),      # line solves
    ((128, 96, 64), 1e-5),      # convolution sweeps of 7, 12 and 8 taps a side
    ((96, 64, 48), 6e-5),       # convolution sweeps of 9, 16 and 11 taps a side
])
def test_periodic_helmholtz_matches_fft(ctx, solver, shape, alpha):
    L = (2 * math.pi, 1.7, 2.3)
    h = [L[a] / shape[a] for a in range(3)]
    g = sa.grid(*shape, *h)
    b = sa.bc_set([PER] * 6)
    rng = np.random.default_rng(5)
    rhs, tor = zip(*[periodic_field(rng, shape, c) for c in range(3)])
    out = zeros_vel(g)
    if solver == "adi":
        ctx.run("of_solve_predictor_adi", C.byref(g), C.byref(b), C.c_double(alpha), *map(sa.ptr, map(dev, rhs)),
                *map(sa.ptr, out), None)
    else:
        it, rel = C.c_int(), C.c_double()
        ctx.run("of_solve_predictor_cg", C.byref(g), C.byref(b), C.c_double(alpha), *map(sa.ptr, map(dev, rhs)),
                *map(sa.ptr, out), C.c_double(1e-13), 200, C.byref(it), C.byref(rel))
        assert rel.value <= 1e-13
    for c in range(3):
        want = periodic_helmholtz_fft(tor[c], h, alpha)
        got = torus(host(out[c]), c)
        assert np.max(np.abs(got - want)) <= 1e-11 * np.max(np.abs(want)), c


# ---------------------------------------------------------------- MGPCG exact discrete recovery
def exact_discrete_poisson(rhs, h, kinds):
    """Solution of lap_h p = rhs (the solve's sign convention) with homogeneous cell-centred
    boundary relations per axis: 'd' Dirichlet (ghost = -interior: sine basis, DST-II), 'n'
    Neumann (ghost = interior: cosine basis, DCT-II; the constant mode is set to zero).
    Eigenvalues (2 - 2 cos(pi k / n)) / h^2, k = 1..n (sine) or 0..n-1 (cosine)."""
    x = rhs.copy()
    lam = np.zeros(rhs.shape)
    for ax in range(3):
        n = rhs.shape[ax]
        kind = kinds[2 - ax]  # kinds are given in (x, y, z) order
        if n == 1:
            continue
        k = np.arange(1, n + 1) if kind == "d" else np.arange(n)
        l1 = (2 - 2 * np.cos(np.pi * k / n)) / h[2 - ax] ** 2
        x = (sfft.dst if kind == "d" else sfft.dct)(x, type=2, axis=ax, norm="ortho")
        shp = [1, 1, 1]
        shp[ax] = n
        lam = lam + l1.reshape(shp)
    with np.errstate(divide="ignore", invalid="ignore"):
        x = np.where(lam > 0, -x / np.where(lam > 0, lam, 1.0), 0.0)
    for ax in range(3):
        if rhs.shape[ax] == 1:
            continue
        kind = kinds[2 - ax]
        x = (sfft.idst if kind == "d" else sfft.idct)(x, type=2, axis=ax, norm="ortho")
    return x


def pbc_for(kinds):
    t = []
    for kd in kinds:
        t += [sa.P_DIRICHLET, sa.P_DIRICHLET] if kd == "d" else [sa.P_NEUMANN, sa.P_NEUMANN]
    return sa.pressure_bc_set(t)


@pytest.mark.parametrize("precision", [sa.FP64, sa.MIXED])
@pytest.mark.parametrize("shape,kinds", [
    ((256, 256, 1), "ddn"),   # Dirichlet walls in x and y (2D)
    ((256, 256, 1), "nnn"),   # pure Neumann (2D): zero-mean solution
    ((64, 64, 64), "ddn"),    # Dirichlet x, y; Neumann z
    ((64, 64, 64), "nnn"),    # pure Neumann (3D)
    ((64, 48, 40), "ndd"),
])
def test_poisson_recovers_exact_discrete_solution(ctx, precision, shape, kinds):
    nx, ny, nz = shape
    h = (1.0 / nx, 1.2 / ny, 0.8 / nz if nz > 1 else 1.0)
    g = sa.grid(nx, ny, nz, *h)
    rng = np.random.default_rng(17)
    rhs = np.zeros(sa.storage_shape(nx, ny, nz, sa.CELL))
    act = rng.standard_normal((nz, ny, nx))
    if "d" not in kinds:
        act -= act.mean()
    rhs[1:-1, 1:-1, 1:-1] = act
    tol = 1e-12
    o = sa.ProjectionOptions(1.0, 1.0, 400, tol)
    sol = zeros(g, sa.CELL)
    d = sa.PoissonDiagnostics()
    ctx.run("of_solve_pressure_poisson", C.byref(g), C.byref(pbc_for(kinds)), C.byref(o), precision,
            sa.ptr(dev(rhs)), sa.ptr(sol), C.byref(d), None, 0)
    assert d.converged == 1 and d.relative_residual <= tol
    assert d.zero_mean_enforced == (1 if "d" not in kinds else 0)
    want = exact_discrete_poisson(act, h, kinds)
    got = host(sol)[1:-1, 1:-1, 1:-1]
    if "d" not in kinds:
        assert abs(got.mean()) <= 1e-12 * np.max(np.abs(want))
    assert np.max(np.abs(got - want)) <= 1e-8 * np.max(np.abs(want))


@pytest.mark.parametrize("case", ["dirichlet", "neumann"])
def test_poisson_manufactured_recovery_reference_cases(ctx, case):
    """tests/test_runtime.cpp:1178-1269: sin(pi x) sin(pi y) with Dirichlet walls at 32^2, and
    cos(2 pi x) cos(2 pi y) (zero mean) with Neumann walls at 24^2; rhs from the operator itself,
    recovery within 1e-8 relative at tolerance 1e-10."""
    n = 32 if case == "dirichlet" else 24
    g = sa.grid(n, n, 1, 1.0 / n, 1.0 / n, 1.0)
    x, y, _ = coords(g, sa.CELL)
    if case == "dirichlet":
        exact = np.sin(np.pi * x) * np.sin(np.pi * y)
        pb = sa.pressure_bc_set([sa.P_DIRICHLET] * 4 + [sa.P_NEUMANN] * 2)
    else:
        exact = np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y)
        exact[1:-1, 1:-1, 1:-1] -= exact[1:-1, 1:-1, 1:-1].mean()
        pb = sa.pressure_bc_set([sa.P_NEUMANN] * 6)
    rhs = zeros(g, sa.CELL)
    ctx.run("of_build_poisson_rhs_from_pressure", C.byref(g), C.byref(pb), sa.ptr(dev(exact)), sa.ptr(rhs))
    sol = zeros(g, sa.CELL)
    d = sa.PoissonDiagnostics()
    o = sa.ProjectionOptions(1.0, 1.0, 300, 1e-10)
    ctx.run("of_solve_pressure_poisson", C.byref(g), C.byref(pb), C.byref(o), sa.FP64, sa.ptr(rhs), sa.ptr(sol),
            C.byref(d), None, 0)
    a, b = host(sol)[1:-1, 1:-1, 1:-1], exact[1:-1, 1:-1, 1:-1]
    assert d.converged == 1 and d.relative_residual <= 1e-10
    assert d.multigrid_levels >= 2 and d.coarse_unknowns > 0
    assert d.pre_smoothing_steps == 2 and d.post_smoothing_steps == 2
    assert np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b * b)) <= 1e-8
    if case == "neumann":
        assert d.zero_mean_enforced == 1 and abs(a.mean()) <= 1e-10


# ---------------------------------------------------------------- TVD boundedness
@pytest.mark.parametrize("shape", [(64, 16, 1), (256, 64, 1), (128, 32, 16)])
def test_tvd_step_update_stays_bounded(ctx, shape):
    """u = 1, v = 0 and a step in x carried by the z (or, in 3D, the y) component; one explicit
    update with dt = 0.4 dx stays within [0, 1] (tests/test_runtime.cpp:1414-1445)."""
    nx, ny, nz = shape
    L = 1.0
    g = sa.grid(nx, ny, nz, L / nx, L / ny, L / nz if nz > 1 else 1.0)
    dt = 0.4 * g.dx
    vel = [np.ones(sa.storage_shape(nx, ny, nz, sa.FACE_X)), np.zeros(sa.storage_shape(nx, ny, nz, sa.FACE_Y)),
           np.zeros(sa.storage_shape(nx, ny, nz, sa.FACE_Z))]
    carried = 2 if nz == 1 else 1
    x, _, _ = coords(g, carried + 1)
    vel[carried] = np.where(np.mod(x, L) < 0.5, 1.0, 0.0)
    out = zeros_vel(g)
    ctx.run("of_compute_advection_term", C.byref(g), 0, *map(sa.ptr, map(dev, vel)), *map(sa.ptr, out))
    upd = vel[carried][1:-1, 1:-1, 1:-1] - dt * host(out[carried])[1:-1, 1:-1, 1:-1]
    assert upd.min() >= -1e-12 and upd.max() <= 1 + 1e-12
    # the update does move the step (the check is not vacuous)
    assert np.max(np.abs(upd - vel[carried][1:-1, 1:-1, 1:-1])) > 0.1


# ---------------------------------------------------------------- projection known answers
@pytest.mark.parametrize("shape", [(16, 12, 1), (256, 256, 1)])
def test_quiescent_projection_takes_no_iteration(ctx, shape):
    nx, ny, nz = shape
    g = sa.grid(nx, ny, nz, 1.0 / nx, 1.0 / ny, 1.0)
    b = sa.bc_set([NS, NS, NS, PV, SYM, SYM], [[0, 0, 0]] * 3 + [[1.0, 0, 0]] + [[0, 0, 0]] * 2)
    o = sa.ProjectionOptions(0.05, 1.0, 2000, 1e-12)
    pred = zeros_vel(g)
    out = zeros_vel(g)
    phi = zeros(g, sa.CELL)
    prhs = zeros(g, sa.CELL)
    d = sa.ProjectionDiagnostics()
    ctx.run("of_project_velocity", C.byref(g), C.byref(b), C.byref(o), sa.FP64, *map(sa.ptr, pred), sa.ptr(phi),
            *map(sa.ptr, out), sa.ptr(prhs), C.byref(d), None, 0)
    ps = d.pressure_solve
    assert ps.converged == 1 and ps.iterations == 0
    assert d.rhs_l2 <= 1e-14 and d.divergence_l2_before <= 1e-14 and d.divergence_l2_after <= 1e-14
    assert abs(d.pressure_mean) <= 1e-14
    # the zero field with the lid's tangential ghosts: interior stays zero
    for t in out:
        assert np.max(np.abs(host(t)[1:-1, 1:-1, 1:-1])) <= 1e-12
    assert np.max(np.abs(host(phi))) <= 1e-14


@pytest.mark.parametrize("shape", [(24, 24, 1), (128, 128, 1), (32, 32, 32)])
def test_gradient_field_projection_recovers_pressure(ctx, shape):
    """Symmetry walls (pure Neumann pressure): predicted = dt grad(p) with p = sin^2(pi x)
    sin^2(pi y) [sin^2(pi z)]; the projection returns p less its mean, zero velocity and zero
    divergence (tests/test_runtime.cpp:1479-1539)."""
    nx, ny, nz = shape
    g = sa.grid(nx, ny, nz, 1.0 / nx, 1.0 / ny, 1.0 / nz if nz > 1 else 1.0)
    types = [SYM] * 6
    b = sa.bc_set(types)
    pb = sa.PressureBCSet()
    call = sa.lib().of_derive_pressure_correction_boundary_conditions
    sa._check(call(C.byref(b), C.byref(pb)))
    dt = 0.025
    x, y, z = coords(g, sa.CELL)
    p = np.sin(np.pi * x) ** 2 * np.sin(np.pi * y) ** 2
    if nz > 1:
        p = p * np.sin(np.pi * z) ** 2
    pd = dev(p)
    ctx.run("of_apply_pressure_boundary_conditions", C.byref(g), C.byref(pb), sa.ptr(pd))
    pred = zeros_vel(g)
    ctx.run("of_compute_gradient", C.byref(g), sa.ptr(pd), *map(sa.ptr, pred))
    for t in pred:
        t.mul_(dt)
    ctx.run("of_apply_velocity_boundary_conditions", C.byref(g), C.byref(b), *map(sa.ptr, pred))
    o = sa.ProjectionOptions(dt, 1.0, 4000, 1e-12)
    out = zeros_vel(g)
    phi = zeros(g, sa.CELL)
    d = sa.ProjectionDiagnostics()
    ctx.run("of_project_velocity", C.byref(g), C.byref(b), C.byref(o), sa.FP64, *map(sa.ptr, pred), sa.ptr(phi),
            *map(sa.ptr, out), None, C.byref(d), None, 0)
    want = host(pd)[1:-1, 1:-1, 1:-1]
    want = want - want.mean()
    got = host(phi)[1:-1, 1:-1, 1:-1]
    assert d.pressure_solve.converged == 1 and d.pressure_solve.zero_mean_enforced == 1
    assert abs(d.pressure_mean) <= 1e-10
    assert np.sqrt(np.mean((got - want) ** 2)) <= 1e-8
    assert max(float(t.abs().max()) for t in out) <= 1e-8
    assert d.divergence_l2_after <= 1e-10
