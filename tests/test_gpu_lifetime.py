"""Native handle lifetime on the device: a stepper or adjoint run keeps its context alive.

of_context_destroy drops the caller's reference only; the native context goes with the last
stepper or adjoint run created on it. A context closed (or garbage-collected) before its
stepper must therefore leave the stepper fully usable, and leave no HIP error pending for
later, unrelated calls.
"""
import gc

import numpy as np
import pytest

import solver_amd as sa

pytestmark = pytest.mark.gpu

N, NU, TOL = 16, 0.05, 1e-12


def run_steps(ctx, steps, close_context_first):
    dt, _ = sa.taylor_green_dt(N, N, N, 0.5, 1.0)
    st = sa.Stepper(ctx, sa.TAYLOR_GREEN, N, N, N, NU, dt, 200, TOL, sa.FP64)
    st.initialize(1)
    if close_context_first:
        ctx.close()
    m = st.run(steps)
    vel, p, _, _, _ = st.download()
    st.close()
    return vel, p, m


def test_stepper_outlives_closed_context():
    ref_ctx = sa.Context(0)
    vel0, p0, m0 = run_steps(ref_ctx, 3, close_context_first=False)
    ref_ctx.close()
    vel1, p1, m1 = run_steps(sa.Context(0), 3, close_context_first=True)
    assert m1.pressure_iterations == m0.pressure_iterations
    for a, b in zip(vel1 + [p1], vel0 + [p0]):
        assert np.array_equal(a, b)


def test_adjoint_outlives_closed_context():
    import torch

    ctx = sa.Context(0)
    dt, _ = sa.taylor_green_dt(N, N, N, 0.5, 1.0)
    adj = sa.Adjoint(ctx, N, N, N, NU, dt, 2, 200, TOL, sa.FP64)
    ctx.close()
    shapes = [sa.storage_shape(N, N, N, loc) for loc in (sa.FACE_X, sa.FACE_Y, sa.FACE_Z)]
    vel = [torch.zeros(s, dtype=torch.float64, device="cuda") for s in shapes]
    ke, _ = adj.forward(vel, None, 2)
    assert ke == 0.0
    gu, _, _ = adj.backward(want_pressure=False, want_forcing=False)
    assert all(float(g.abs().max()) == 0.0 for g in gu)
    adj.close()


def _garbage_cycle():
    ctx = sa.Context(0)
    dt, _ = sa.taylor_green_dt(N, N, N, 0.5, 1.0)
    st = sa.Stepper(ctx, sa.TAYLOR_GREEN, N, N, N, NU, dt, 1, TOL, sa.FP64)
    st.initialize(1)
    try:
        st.run(1)  # one pressure iteration at tolerance 1e-12: the reference's non-convergence error
    except sa.SolverError as e:
        err = e
    # err -> traceback -> this frame -> {ctx, st}: only the cycle collector frees them
    return err


def test_collected_cycle_leaves_no_pending_error():
    assert "did not converge" in str(_garbage_cycle())
    gc.collect()
    # a HIP error left by the collected handles would fail this run's launch checks
    vel, p, m = run_steps(sa.Context(0), 2, close_context_first=False)
    assert m.divergence_l2 <= 1e-10
    assert all(np.isfinite(a).all() for a in vel + [p])
