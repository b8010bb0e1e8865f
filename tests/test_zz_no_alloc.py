"""The allocation-free pressure solve (of_solve_pressure_poisson_noalloc) against the
allocating one (of_solve_pressure_poisson), through the C ABI.

Both run the same kernels at the same launch shapes, so the results must agree bit for bit
(solution, iteration count, residuals). Two calls in a row, the second on a smaller grid with
other boundary conditions, check that the scratch block keeps nothing from the previous call
that matters; the sizes cover 3D periodic grids (fused multigrid levels and the full-row CG
kernel) and wall/Dirichlet grids with a coarsest system in global memory.
"""
import ctypes as C

import numpy as np
import pytest

import solver_amd as sa
from tests.gpu_util import dev, host, zeros

pytestmark = pytest.mark.gpu

CASES = [
    # (shape, pressure BC types, face values, face gradients)
    ((64, 64, 64), [sa.P_PERIODIC] * 6, None, None),
    ((40, 48, 36), [sa.P_PERIODIC] * 6, None, None),
    ((18, 12, 8), [2, 2, 0, 0, 1, 1], [0.0, 0.0, 0.2, -0.1, 0.3, 0.0], [0.0, 0.0, 0.1, 0.0, 0.0, -0.2]),
    ((24, 20, 1), [1, 1, 0, 0, 0, 0], [0.4, -0.2, 0.0, 0.0, 0.0, 0.0], [0.0, 0.0, 0.3, 0.1, 0.0, 0.0]),
]


@pytest.fixture(scope="module")
def ctx():
    c = sa.Context(0)
    yield c
    c.close()


def inputs(shape, seed):
    nx, ny, nz = shape
    h = (1.0 / nx, 1.3 / ny, 0.9 / nz if nz > 1 else 1.0)
    g = sa.grid(nx, ny, nz, *h)
    rng = np.random.default_rng(seed)
    rhs = rng.standard_normal(sa.storage_shape(nx, ny, nz, sa.CELL))
    warm = 0.1 * rng.standard_normal(sa.storage_shape(nx, ny, nz, sa.CELL))
    return g, rhs, warm


def solve_allocating(ctx, g, pb, o, precision, rhs, warm):
    p = dev(warm)
    d = sa.PoissonDiagnostics()
    r = dev(rhs)
    ctx.run("of_solve_pressure_poisson", C.byref(g), C.byref(pb), C.byref(o), precision, sa.ptr(r), sa.ptr(p),
            C.byref(d), None, 0)
    return host(p), d


def solve_noalloc(g, pb, o, precision, rhs, warm):
    import torch

    p = dev(warm)
    d = sa.PoissonDiagnostics()
    r = dev(rhs)
    torch.cuda.synchronize()
    sa.call("of_solve_pressure_poisson_noalloc", 0, None, C.byref(g), C.byref(pb), C.byref(o), precision,
            sa.ptr(r), sa.ptr(p), C.byref(d))
    torch.cuda.synchronize()
    return host(p), d


@pytest.mark.parametrize("precision", [sa.FP64, sa.MIXED])
@pytest.mark.parametrize("first,second", [(0, 2), (1, 3), (0, 1)])
def test_noalloc_matches_allocating_solve_twice(ctx, precision, first, second):
    for k, case in enumerate((CASES[first], CASES[second])):
        shape, types, vals, grads = case
        print(f"noalloc case {shape} bc {types} precision {precision}", flush=True)
        g, rhs, warm = inputs(shape, 100 * first + 10 * second + k)
        pb = sa.pressure_bc_set(types, vals, grads)
        o = sa.ProjectionOptions(1.0, 1.0, 400, 1e-11)
        pa, da = solve_allocating(ctx, g, pb, o, precision, rhs, warm)
        pn, dn = solve_noalloc(g, pb, o, precision, rhs, warm)
        assert da.converged == 1 and dn.converged == 1
        assert dn.iterations == da.iterations
        assert dn.relative_residual == da.relative_residual
        assert dn.initial_residual_l2 == da.initial_residual_l2
        assert dn.multigrid_levels == da.multigrid_levels and dn.coarse_unknowns == da.coarse_unknowns
        assert np.array_equal(pn, pa), float(np.max(np.abs(pn - pa)))
    assert np.prod(CASES[second][0]) < np.prod(CASES[first][0])


def test_noalloc_rejects_a_grid_beyond_its_scratch():
    g = sa.grid(128, 64, 64, 0.1, 0.1, 0.1)
    pb = sa.pressure_bc_set([sa.P_PERIODIC] * 6)
    o = sa.ProjectionOptions(1.0, 1.0, 10, 1e-8)
    dummy = zeros(sa.grid(2, 2, 2, 1.0, 1.0, 1.0), sa.CELL)
    with pytest.raises(sa.InvalidArgument, match="at most 262144 cells"):
        sa.call("of_solve_pressure_poisson_noalloc", 0, None, C.byref(g), C.byref(pb), C.byref(o), sa.FP64,
                sa.ptr(dummy), sa.ptr(dummy), None)
