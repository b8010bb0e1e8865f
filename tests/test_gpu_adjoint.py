"""GPU checks of the discrete adjoint and tangent-linear model.

* Dot-product tests <J x, y> = <x, J^T y> for every transposed stage through the C ABI
  (velocity and pressure ghost fills, gradient, divergence, ADI predictor, linearised
  advection) on random fields, to rounding.
* The device advection tangent against the oracle's dual-number tangent.
* Whole runs (checkpointed): the device tangent of the final kinetic energy against the
  oracle's tangent-linear model (itself pinned by central differences on the CPU); the
  adjoint gradient against the device tangent in the same directions (velocity and forcing),
  for several checkpoint intervals and for an arbitrary seed; the adjoint gradient against
  central finite differences of the device forward run.
* torch.autograd: the Functions' gradients equal the adjoint driver's.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import solver_amd as sa
from oracle import ofref
from solver_amd import inputs
from tests.gpu_util import dev, host

pytestmark = pytest.mark.gpu

LOCS = (sa.FACE_X, sa.FACE_Y, sa.FACE_Z)


@pytest.fixture(scope="module")
def ctx():
    c = sa.Context(0)
    yield c
    c.close()


def vdot(a, b):
    return float(sum(np.sum(np.asarray(x) * np.asarray(y)) for x, y in zip(a, b)))


def rand_fields(rng, g, locs):
    return [rng.standard_normal(sa.storage_shape(g.nx, g.ny, g.nz, loc)) for loc in locs]


def periodic_grid(nx, ny, nz):
    L = 2 * math.pi
    return sa.grid(nx, ny, nz, L / nx, L / ny, L / nz)


def ptrs(ts):
    return [sa.ptr(t) for t in ts]


# ------------------------------------------------------------------ per-stage dot products
def test_velocity_and_pressure_bc_adjoint(ctx):
    g = periodic_grid(7, 6, 5)
    b = sa.taylor_green_bcs(5)
    pb = sa.pressure_bc_set([sa.P_PERIODIC] * 6)
    rng = np.random.default_rng(0)
    x, y = rand_fields(rng, g, LOCS), rand_fields(rng, g, LOCS)
    bx, bty = [dev(a) for a in x], [dev(a) for a in y]
    ctx.run("of_apply_velocity_boundary_conditions", C.byref(g), C.byref(b), *ptrs(bx))
    ctx.run("of_apply_velocity_boundary_conditions_adjoint", C.byref(g), C.byref(b), *ptrs(bty))
    lhs, rhs = vdot(map(host, bx), y), vdot(x, map(host, bty))
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)
    p, q = rand_fields(rng, g, [sa.CELL])[0], rand_fields(rng, g, [sa.CELL])[0]
    bp, btq = dev(p), dev(q)
    ctx.run("of_apply_pressure_boundary_conditions", C.byref(g), C.byref(pb), sa.ptr(bp))
    ctx.run("of_apply_pressure_boundary_conditions_adjoint", C.byref(g), C.byref(pb), sa.ptr(btq))
    lhs, rhs = float(np.sum(host(bp) * q)), float(np.sum(p * host(btq)))
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)
    # non-periodic boundaries are rejected
    with pytest.raises(sa.InvalidArgument, match="periodic faces only"):
        ctx.run("of_apply_velocity_boundary_conditions_adjoint", C.byref(g), C.byref(sa.bc_set([sa.NO_SLIP] * 6)),
                *ptrs(bty))


def active_mask(g, loc):
    m = np.zeros(sa.storage_shape(g.nx, g.ny, g.nz, loc), dtype=bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m


def test_gradient_and_divergence_adjoint(ctx):
    g = periodic_grid(6, 5, 7)
    rng = np.random.default_rng(1)
    p = rng.standard_normal(sa.storage_shape(6, 5, 7, sa.CELL))
    gy = [a * active_mask(g, loc) for a, loc in zip(rand_fields(rng, g, LOCS), LOCS)]
    out = [dev(np.zeros_like(a)) for a in gy]
    ctx.run("of_compute_gradient", C.byref(g), sa.ptr(dev(p)), *ptrs(out))
    pt = dev(np.zeros_like(p))
    ctx.run("of_compute_gradient_adjoint", C.byref(g), *ptrs([dev(a) for a in gy]), sa.ptr(pt))
    lhs = vdot([host(o) * active_mask(g, loc) for o, loc in zip(out, LOCS)], gy)
    rhs = float(np.sum(p * host(pt)))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    u = rand_fields(rng, g, LOCS)
    d = rng.standard_normal(p.shape) * active_mask(g, sa.CELL)
    dv = dev(np.zeros_like(p))
    ctx.run("of_compute_divergence", C.byref(g), *ptrs([dev(a) for a in u]), sa.ptr(dv))
    ut = [dev(np.zeros_like(a)) for a in u]
    ctx.run("of_compute_divergence_adjoint", C.byref(g), sa.ptr(dev(d)), *ptrs(ut))
    lhs = float(np.sum(host(dv) * active_mask(g, sa.CELL) * d))
    rhs = vdot(u, map(host, ut))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


# (40, 18, 34): lines of several 16-element segments of the two-pass line solver
@pytest.mark.parametrize("alpha,shape", [(0.0, (9, 8, 6)), (0.02, (9, 8, 6)), (0.7, (9, 8, 6)), (0.7, (40, 18, 34))])
def test_predictor_adi_adjoint(ctx, alpha, shape):
    g = periodic_grid(*shape)
    b = sa.taylor_green_bcs(shape[2])
    rng = np.random.default_rng(2)
    r, y = rand_fields(rng, g, LOCS), rand_fields(rng, g, LOCS)
    out = [dev(np.zeros_like(a)) for a in r]
    ctx.run("of_solve_predictor_adi", C.byref(g), C.byref(b), C.c_double(alpha), *ptrs([dev(a) for a in r]),
            *ptrs(out), None)
    rt = [dev(np.zeros_like(a)) for a in y]
    ctx.run("of_solve_predictor_adi_adjoint", C.byref(g), C.byref(b), C.c_double(alpha), *ptrs([dev(a) for a in y]),
            *ptrs(rt))
    lhs, rhs = vdot(map(host, out), y), vdot(r, map(host, rt))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)


@pytest.mark.parametrize("scheme", [0, 1, 2])
def test_advection_tangent_adjoint_and_oracle(ctx, scheme):
    g = periodic_grid(7, 6, 5)
    rng = np.random.default_rng(3 + scheme)
    u, du, a = rand_fields(rng, g, LOCS), rand_fields(rng, g, LOCS), rand_fields(rng, g, LOCS)
    tl = [dev(np.zeros_like(x)) for x in u]
    ctx.run("of_compute_advection_tangent", C.byref(g), scheme, *ptrs([dev(x) for x in u]),
            *ptrs([dev(x) for x in du]), *ptrs(tl))
    adj = [dev(np.zeros_like(x)) for x in u]
    ctx.run("of_compute_advection_adjoint", C.byref(g), scheme, *ptrs([dev(x) for x in u]),
            *ptrs([dev(x) for x in a]), *ptrs(adj))
    # the adjoint is a gather with a fixed summation order: a second call is bitwise equal
    adj2 = [dev(np.zeros_like(x)) for x in u]
    ctx.run("of_compute_advection_adjoint", C.byref(g), scheme, *ptrs([dev(x) for x in u]),
            *ptrs([dev(x) for x in a]), *ptrs(adj2))
    for x, y in zip(adj, adj2):
        assert np.array_equal(host(x), host(y))
    lhs = vdot(map(host, tl), [x * active_mask(g, loc) for x, loc in zip(a, LOCS)])
    rhs = vdot(du, map(host, adj))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)
    og = ofref.make_grid(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz)
    ref = ofref.advection_tangent(og, u, du, scheme)
    for c in range(3):
        err = np.max(np.abs(host(tl[c]) - ref[c])) / np.max(np.abs(ref[c]))
        assert err <= 1e-12, (c, err)


# ------------------------------------------------------------------ whole runs
SEED = 20261203


def tg_setup(n, steps, nu=0.05, amp=0.05, tol=1e-13):
    probe = ofref.tg_config(n, n, n, viscosity=nu, cfl_limit=0.5, final_time=1.0)
    upper = ofref.tg_dt(probe)
    c = ofref.tg_config(n, n, n, viscosity=nu, cfl_limit=0.5, final_time=steps * upper * 0.999,
                        max_iterations=400, tolerance=tol)
    s = ofref.tg_init(c, convention=1)
    vel = [a + b for a, b in zip(s.vel, inputs.perturbation(n, SEED, amp))]
    return c, ofref.tg_dt(c), vel, s.p


def make_adjoint(ctx, c, dt, every, precision=sa.FP64, scheme=0):
    return sa.Adjoint(ctx, c.nx, c.ny, c.nz, c.viscosity, dt, every, c.max_iterations, c.tolerance, precision, scheme)


def test_forward_matches_oracle_with_forcing(ctx):
    n, steps = 12, 4
    c, dt, vel, p = tg_setup(n, steps)
    force = inputs.forcing(n, 7, 0.3)
    adj = make_adjoint(ctx, c, dt, 2)
    adj.set_forcing([dev(a) for a in force])
    ke, m = adj.forward([dev(a) for a in vel], dev(p), steps)
    u_ref, p_ref, ke_ref, _, _ = ofref.tg_run_tl(c, steps, vel, p, force=force)
    assert abs(ke - ke_ref) <= 1e-12 * ke_ref
    fin = adj.final_velocity()
    for a, b in zip(fin, u_ref):
        assert np.max(np.abs(host(a) - b)) <= 1e-10 * np.max(np.abs(b))
    assert m.step == steps and m.divergence_linf <= 1e-10


@pytest.mark.parametrize("every", [1, 2, 5])
def test_tangent_and_adjoint_agree_with_oracle(ctx, every):
    n, steps = 12, 5
    c, dt, vel, p = tg_setup(n, steps)
    force = inputs.forcing(n, 7, 0.3)
    dvel = inputs.perturbation(n, SEED + 1, 1.0)
    dforce = inputs.perturbation(n, SEED + 2, 1.0)
    adj = make_adjoint(ctx, c, dt, every)
    adj.set_forcing([dev(a) for a in force])
    adj.forward([dev(a) for a in vel], dev(p), steps)
    _, dke_u = adj.tangent(dvel=[dev(a) for a in dvel])
    _, dke_f = adj.tangent(dforce=[dev(a) for a in dforce])
    _, _, _, _, ref_u = ofref.tg_run_tl(c, steps, vel, p, force=force, dvel=dvel)
    _, _, _, _, ref_f = ofref.tg_run_tl(c, steps, vel, p, force=force, dvel=[np.zeros_like(a) for a in vel],
                                        dforce=dforce)
    assert abs(dke_u - ref_u) <= 1e-9 * abs(ref_u), (dke_u, ref_u)
    assert abs(dke_f - ref_f) <= 1e-9 * abs(ref_f), (dke_f, ref_f)
    gu, gp, gf = adj.backward()
    adj_u = vdot(map(host, gu), dvel)
    adj_f = vdot(map(host, gf), dforce)
    assert abs(adj_u - dke_u) <= 1e-10 * abs(dke_u), (adj_u, dke_u)
    assert abs(adj_f - dke_f) <= 1e-10 * abs(dke_f), (adj_f, dke_f)
    # gradients are confined to active entries; the pressure gradient vanishes on a periodic box
    for t, loc in zip(gu, LOCS):
        assert not np.any(host(t)[~active_mask(c, loc)])
    assert np.max(np.abs(host(gp))) <= 1e-10 * max(np.max(np.abs(host(t))) for t in gu)


def test_adjoint_with_arbitrary_seed(ctx):
    n, steps = 10, 4
    c, dt, vel, p = tg_setup(n, steps)
    adj = make_adjoint(ctx, c, dt, 3)
    adj.forward([dev(a) for a in vel], dev(p), steps)
    rng = np.random.default_rng(5)
    seed = rand_fields(rng, c, LOCS)
    dvel = inputs.perturbation(n, SEED + 3, 1.0)
    du_n, _ = adj.tangent(dvel=[dev(a) for a in dvel])
    gu, _, _ = adj.backward(seed=[dev(a) for a in seed], want_pressure=False, want_forcing=False)
    lhs = vdot(map(host, du_n), seed)
    rhs = vdot(dvel, map(host, gu))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs), (lhs, rhs)


@pytest.mark.parametrize("precision", [sa.FP64, sa.MIXED])
def test_adjoint_gradient_matches_central_differences(ctx, precision):
    n, steps = 16, 6
    c, dt, vel, p = tg_setup(n, steps, nu=0.02, amp=0.05, tol=1e-12)
    adj = make_adjoint(ctx, c, dt, 3, precision)
    ke0, _ = adj.forward([dev(a) for a in vel], dev(p), steps)
    gu, _, _ = adj.backward(want_pressure=False, want_forcing=False)
    g = [host(t) for t in gu]
    for d in range(1, 3):
        dvel = inputs.perturbation(n, SEED + d, 1.0)
        pred = vdot(g, dvel)
        errs = []
        for eps in (1e-5, 1e-6, 1e-7):
            kp, _ = adj.forward([dev(a + eps * b) for a, b in zip(vel, dvel)], dev(p), steps)
            km, _ = adj.forward([dev(a - eps * b) for a, b in zip(vel, dvel)], dev(p), steps)
            errs.append(abs((kp - km) / (2 * eps) - pred) / abs(pred))
        assert min(errs) <= 1e-6, (d, errs)


def test_autograd_functions(ctx):
    torch = pytest.importorskip("torch")
    from solver_amd.autograd import rollout, rollout_kinetic_energy

    n, steps = 10, 3
    c, dt, vel, p = tg_setup(n, steps)
    force = inputs.forcing(n, 9, 0.2)
    adj = make_adjoint(ctx, c, dt, 2)
    u0 = [dev(a).requires_grad_() for a in vel]
    f = [dev(a).requires_grad_() for a in force]
    ke = rollout_kinetic_energy(adj, steps, *u0, dev(p), forcing=f)
    ke.backward()
    gu, _, gf = adj.backward(want_pressure=False)
    # two reverse sweeps are bitwise equal (every adjoint kernel sums in a fixed order)
    for a, b in zip(list(u0) + list(f), list(gu) + list(gf)):
        assert torch.equal(a.grad, b)
    # rollout: vector-Jacobian product with an arbitrary cotangent
    adj2 = make_adjoint(ctx, c, dt, 2)
    u1 = [dev(a).requires_grad_() for a in vel]
    out = rollout(adj2, steps, *u1, dev(p))
    rng = np.random.default_rng(8)
    w = [dev(rng.standard_normal(t.shape)) for t in out]
    grads = torch.autograd.grad(out, u1, grad_outputs=w)
    dvel = inputs.perturbation(n, SEED + 4, 1.0)
    du_n, _ = adj2.tangent(dvel=[dev(a) for a in dvel])
    lhs = vdot([host(t) for t in du_n], [host(t) for t in w])
    rhs = vdot(dvel, [host(t) for t in grads])
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_adjoint_rejects_unsupported_cases(ctx):
    with pytest.raises(sa.InvalidArgument, match="three-dimensional periodic Taylor-Green"):
        sa.Adjoint(ctx, 16, 16, 1, 0.01, 0.01, 2)
    with pytest.raises(sa.InvalidArgument, match="checkpoint interval"):
        sa.Adjoint(ctx, 8, 8, 8, 0.01, 0.01, 0)
    adj = sa.Adjoint(ctx, 8, 8, 8, 0.01, 0.01, 2)
    with pytest.raises(sa.InvalidArgument, match="forward pass"):
        adj.backward()


def test_forced_mixed_gradient_wrt_forcing(ctx):
    """The C5 workload's shape at a small size: mixed precision (FP32 V-cycle inside FP64 CG),
    Re 1600, the seeded forcing f = (A sin(kx x) sin(ky y), 0, 0), checkpoints every 2 of 6
    steps; the gradient of the final kinetic energy w.r.t. f against the device tangent and
    against central differences of the forward run along seeded forcing directions."""
    n, steps = 16, 6
    c, dt, vel, p = tg_setup(n, steps, nu=1 / 1600, amp=0.05, tol=1e-12)
    base = inputs.forcing(n, SEED, 1e-2)
    adj = make_adjoint(ctx, c, dt, 2, sa.MIXED)
    adj.set_forcing([dev(a) for a in base])
    u0 = [dev(a) for a in vel]
    adj.forward(u0, dev(p), steps)
    _, _, gf = adj.backward(want_pressure=False, want_forcing=True)
    g = [host(t) for t in gf]
    assert np.max(np.abs(g[0])) > 0
    for d in range(1, 3):
        df = inputs.forcing(n, SEED + d, 1.0)
        pred = vdot(g, df)
        _, dke = adj.tangent(dforce=[dev(a) for a in df])
        assert abs(dke - pred) <= 1e-9 * abs(pred), (d, dke, pred)
        errs = []
        for eps in (1e-4, 1e-5, 1e-6):
            ks = []
            for sgn in (1.0, -1.0):
                adj.set_forcing([dev(b + sgn * eps * x) for b, x in zip(base, df)])
                ks.append(adj.forward(u0, dev(p), steps)[0])
            errs.append(abs((ks[0] - ks[1]) / (2 * eps) - pred) / abs(pred))
        adj.set_forcing([dev(a) for a in base])
        adj.forward(u0, dev(p), steps)
        assert min(errs) <= 1e-6, (d, errs)
    adj.close()


def test_overlapped_reverse_sweep_matches_serial(ctx):
    """The reverse sweep recomputes window w - 1 on a second stepper, stream and host thread while
    the adjoint steps of window w run; the gradients (velocity, pressure, forcing) are bitwise
    those of the serial sweep, over several windows and a short last one."""
    n, steps = 12, 7
    c, dt, vel, p = tg_setup(n, steps)
    force = inputs.forcing(n, 5, 0.1)
    grads = []
    for flag in ("0", "1"):
        os.environ["SOLVER_AMD_ADJOINT_OVERLAP"] = flag
        try:
            adj = make_adjoint(ctx, c, dt, 2)
            adj.set_forcing([dev(a) for a in force])
            adj.forward([dev(a) for a in vel], dev(p), steps)
            gu, gp, gf = adj.backward(want_pressure=True, want_forcing=True)
            grads.append([host(t) for t in list(gu) + [gp] + list(gf)])
            adj.close()
        finally:
            os.environ.pop("SOLVER_AMD_ADJOINT_OVERLAP", None)
    for a, b in zip(*grads):
        assert np.array_equal(a, b)
