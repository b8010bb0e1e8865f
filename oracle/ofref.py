"""ctypes wrapper of oracle/libofref.so, the FP64 CPU restatement of the reference path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.

Fields are numpy float64 arrays of shape (sz, sy, sx) holding the full ghosted storage of the
reference layout (core/field_layout.hpp:69-290): flat index i + sx*(j + sy*k).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libofref.so")

CELL, FACE_X, FACE_Y, FACE_Z = 0, 1, 2, 3
NO_SLIP, PRESCRIBED, SYMMETRY, FIXED_PRESSURE, PERIODIC = 0, 1, 2, 3, 4
NEUMANN, DIRICHLET, P_PERIODIC = 0, 1, 2


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double)]


class BC(C.Structure):
    _fields_ = [("type", C.c_int), ("velocity", C.c_double * 3), ("pressure", C.c_double),
                ("pressure_gradient", C.c_double)]


class BCSet(C.Structure):
    _fields_ = [("f", BC * 6)]


class PBC(C.Structure):
    _fields_ = [("type", C.c_int), ("value", C.c_double), ("gradient", C.c_double)]


class PBCSet(C.Structure):
    _fields_ = [("f", PBC * 6)]


class ProjOpts(C.Structure):
    _fields_ = [("dt", C.c_double), ("density", C.c_double), ("max_iterations", C.c_int),
                ("tolerance", C.c_double)]


class PoissonDiag(C.Structure):
    _fields_ = [("iterations", C.c_int), ("initial_residual_l2", C.c_double),
                ("final_residual_l2", C.c_double), ("relative_residual", C.c_double),
                ("removed_rhs_mean", C.c_double), ("multigrid_levels", C.c_int),
                ("coarse_unknowns", C.c_int), ("pre_smoothing_steps", C.c_int),
                ("post_smoothing_steps", C.c_int), ("converged", C.c_int),
                ("zero_mean_enforced", C.c_int)]


class ProjDiag(C.Structure):
    _fields_ = [("rhs_l2", C.c_double), ("divergence_l2_before", C.c_double),
                ("divergence_l2_after", C.c_double), ("pressure_mean", C.c_double),
                ("poisson", PoissonDiag)]


class StepMetrics(C.Structure):
    _fields_ = [("step", C.c_int), ("time", C.c_double), ("dt", C.c_double),
                ("max_cfl", C.c_double), ("max_velocity_change", C.c_double),
                ("divergence_l2", C.c_double), ("max_divergence_l2", C.c_double),
                ("pressure_iterations", C.c_int), ("pressure_relative_residual", C.c_double),
                ("divergence_linf", C.c_double)]


class TGConfig(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("viscosity", C.c_double),
                ("cfl_limit", C.c_double), ("final_time", C.c_double),
                ("max_iterations", C.c_int), ("tolerance", C.c_double)]


class ChannelConfig(C.Structure):
    _fields_ = [("poiseuille", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("viscosity", C.c_double),
                ("top_velocity", C.c_double), ("pressure_drop", C.c_double), ("cfl_limit", C.c_double),
                ("steps", C.c_int), ("max_iterations", C.c_int), ("tolerance", C.c_double)]


class CavityConfig(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("reynolds", C.c_double),
                ("lid_velocity", C.c_double), ("cfl_limit", C.c_double),
                ("max_steps", C.c_int), ("min_steps", C.c_int),
                ("steady_tolerance", C.c_double), ("max_iterations", C.c_int),
                ("tolerance", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle ofref)")
        _lib = C.CDLL(LIB_PATH)
        _lib.ofref_last_error.restype = C.c_char_p
    return _lib


class OracleError(RuntimeError):
    pass


def _call(name, *args):
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        raise OracleError(lib().ofref_last_error().decode())


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def storage_shape(nx, ny, nz, loc):
    a = [nx, ny, nz]
    if loc != CELL:
        a[loc - 1] += 1
    return (a[2] + 2, a[1] + 2, a[0] + 2)


def make_grid(nx, ny, nz, dx, dy, dz) -> Grid:
    return Grid(nx, ny, nz, dx, dy, dz)


def zeros(g: Grid, loc):
    return np.zeros(storage_shape(g.nx, g.ny, g.nz, loc))


def zeros_vel(g: Grid):
    return [zeros(g, FACE_X), zeros(g, FACE_Y), zeros(g, FACE_Z)]


def bcset(types, velocity=None, pressure=None) -> BCSet:
    b = BCSet()
    for f in range(6):
        b.f[f].type = int(types[f])
        for c in range(3):
            b.f[f].velocity[c] = 0.0 if velocity is None else float(velocity[f][c])
        b.f[f].pressure = 0.0 if pressure is None else float(pressure[f])
        b.f[f].pressure_gradient = 0.0
    return b


def pbcset(types, values=None, gradients=None) -> PBCSet:
    b = PBCSet()
    for f in range(6):
        b.f[f].type = int(types[f])
        b.f[f].value = 0.0 if values is None else float(values[f])
        b.f[f].gradient = 0.0 if gradients is None else float(gradients[f])
    return b


# ------------------------------------------------------------------ operations
def velocity_bc(g, b, vel):
    _call("ofref_velocity_bc", C.byref(g), C.byref(b), *map(_p, vel))


def pressure_bc(g, b, p):
    _call("ofref_pressure_bc", C.byref(g), C.byref(b), _p(p))


def total_pressure_bc(g, b, src, p):
    _call("ofref_total_pressure_bc", C.byref(g), C.byref(b), *map(_p, src), _p(p))


def derive_pressure_bc(b) -> PBCSet:
    out = PBCSet()
    _call("ofref_derive_pressure_bc", C.byref(b), C.byref(out))
    return out


def gradient(g, p):
    out = zeros_vel(g)
    _call("ofref_gradient", C.byref(g), _p(p), *map(_p, out))
    return out


def divergence(g, vel):
    out = zeros(g, CELL)
    _call("ofref_divergence", C.byref(g), *map(_p, vel), _p(out))
    return out


def laplacian(g, loc, field):
    out = np.zeros_like(field)
    _call("ofref_laplacian", C.byref(g), C.c_int(loc), _p(field), _p(out))
    return out


def advection(g, vel, scheme=0):
    out = zeros_vel(g)
    _call("ofref_advection", C.byref(g), C.c_int(scheme), *map(_p, vel), *map(_p, out))
    return out


def advection_tangent(g, vel, dvel, scheme=0):
    """Linearised advection at vel along dvel, branches frozen at vel (dual numbers)."""
    out = zeros_vel(g)
    vel = [np.ascontiguousarray(a, dtype=np.float64).copy() for a in vel]
    dvel = [np.ascontiguousarray(a, dtype=np.float64).copy() for a in dvel]
    _call("ofref_advection_tangent", C.byref(g), C.c_int(scheme), *map(_p, vel), *map(_p, dvel), *map(_p, out))
    return out


def diffusion(g, nu, vel):
    out = zeros_vel(g)
    _call("ofref_diffusion", C.byref(g), C.c_double(nu), *map(_p, vel), *map(_p, out))
    return out


def cfl(g, vel, dt):
    out = np.zeros(5)
    _call("ofref_cfl", C.byref(g), *map(_p, vel), C.c_double(dt), _p(out))
    return dict(zip(["dt", "max_cfl", "max_u", "max_v", "max_w"], out.tolist()))


def factorized_correction(g, vel, alpha, cavity=False):
    out = zeros_vel(g)
    _call("ofref_factorized_correction", C.byref(g), C.c_int(int(cavity)), C.c_double(alpha),
          *map(_p, vel), *map(_p, out))
    return out


def predictor_rhs(g, cur, p, prev, nu, dt, cavity=False, scheme=0):
    """cur must carry velocity BC ghosts; prev=None selects the first-step form."""
    adv, rhs = zeros_vel(g), zeros_vel(g)
    pv = prev if prev is not None else zeros_vel(g)
    _call("ofref_predictor_rhs", C.byref(g), C.c_int(int(cavity)), C.c_int(scheme),
          C.c_double(nu), C.c_double(dt), C.c_int(prev is not None), *map(_p, cur), _p(p),
          *map(_p, pv), *map(_p, adv), *map(_p, rhs))
    return adv, rhs


def predictor_adi(g, b, alpha, rhs):
    out = zeros_vel(g)
    ls = C.c_int(0)
    _call("ofref_predictor_adi", C.byref(g), C.byref(b), C.c_double(alpha), *map(_p, rhs),
          *map(_p, out), C.byref(ls))
    return out, ls.value


def opts(dt=1.0, density=1.0, max_iterations=4000, tolerance=1e-12) -> ProjOpts:
    return ProjOpts(dt, density, max_iterations, tolerance)


def build_pressure_rhs(g, b, o, vel):
    out = zeros(g, CELL)
    _call("ofref_build_pressure_rhs", C.byref(g), C.byref(b), C.byref(o), *map(_p, vel), _p(out))
    return out


def correct_velocity(g, b, o, vel, phi):
    out = zeros_vel(g)
    _call("ofref_correct_velocity", C.byref(g), C.byref(b), C.byref(o), *map(_p, vel), _p(phi),
          *map(_p, out))
    return out


def poisson(g, b, o, rhs, p0=None, history_capacity=4096):
    p = zeros(g, CELL) if p0 is None else p0.copy()
    d = PoissonDiag()
    hist = np.zeros(history_capacity)
    n = C.c_int(0)
    _call("ofref_poisson", C.byref(g), C.byref(b), C.byref(o), _p(rhs), _p(p), C.byref(d),
          _p(hist), C.c_int(history_capacity), C.byref(n))
    return p, d, hist[: min(n.value, history_capacity)].copy()


def poisson_apply(g, b, p):
    out = zeros(g, CELL)
    _call("ofref_poisson_apply", C.byref(g), C.byref(b), _p(p), _p(out))
    return out


def poisson_rhs_from_pressure(g, b, p):
    out = zeros(g, CELL)
    _call("ofref_poisson_rhs_from_pressure", C.byref(g), C.byref(b), _p(p), _p(out))
    return out


def project_velocity(g, b, o, vel, phi0=None):
    phi = zeros(g, CELL) if phi0 is None else phi0.copy()
    corr = zeros_vel(g)
    prhs = zeros(g, CELL)
    d = ProjDiag()
    _call("ofref_project_velocity", C.byref(g), C.byref(b), C.byref(o), *map(_p, vel), _p(phi),
          *map(_p, corr), _p(prhs), C.byref(d))
    return phi, corr, prhs, d


def tg_config(nx, ny, nz, viscosity=0.01, cfl_limit=0.5, final_time=0.5, max_iterations=200,
              tolerance=1e-10) -> TGConfig:
    return TGConfig(nx, ny, nz, viscosity, cfl_limit, final_time, max_iterations, tolerance)


def tg_grid(c: TGConfig) -> Grid:
    L = 2.0 * np.pi
    return Grid(c.nx, c.ny, c.nz, L / c.nx, L / c.ny, L / c.nz if c.nz > 1 else 1.0)


def tg_dt(c: TGConfig) -> float:
    dt = C.c_double(0)
    _call("ofref_tg_dt", C.byref(c), C.byref(dt))
    return dt.value


class State:
    """(u, v, w, p_total, A_prev x3, has_prev, metrics) as in solver/taylor_green.hpp:44-57."""

    def __init__(self, g: Grid):
        self.vel = zeros_vel(g)
        self.p = zeros(g, CELL)
        self.adv = zeros_vel(g)
        self.has_prev = C.c_int(0)
        self.metrics = StepMetrics()

    def copy(self):
        s = State.__new__(State)
        s.vel = [a.copy() for a in self.vel]
        s.p = self.p.copy()
        s.adv = [a.copy() for a in self.adv]
        s.has_prev = C.c_int(self.has_prev.value)
        s.metrics = StepMetrics()
        C.pointer(s.metrics)[0] = self.metrics
        return s


def tg_init(c: TGConfig, convention=0) -> State:
    g = tg_grid(c)
    s = State(g)
    _call("ofref_tg_init", C.byref(c), C.c_int(convention), *map(_p, s.vel), _p(s.p))
    s.metrics.dt = tg_dt(c)
    return s


def tg_steps(c: TGConfig, s: State, steps: int, want_linf=False):
    _call("ofref_tg_steps", C.byref(c), C.c_int(steps), C.c_int(int(want_linf)), *map(_p, s.vel),
          _p(s.p), *map(_p, s.adv), C.byref(s.has_prev), C.byref(s.metrics))
    return s


def tg_run_tl(c: TGConfig, steps, vel, p=None, force=None, dvel=None, dp=None, dforce=None, scheme=0):
    """Runs `steps` steps from (vel, p) with an optional body force, and the tangent-linear model
    along (dvel, dp, dforce) when dvel is given. Only active entries of the inputs are read.
    Returns (final velocity, final pressure, final kinetic energy, final velocity perturbation
    or None, directional derivative of the final kinetic energy or None)."""
    g = tg_grid(c)

    def own(a, loc):
        return zeros(g, loc) if a is None else np.ascontiguousarray(a, dtype=np.float64).copy()

    u = [own(a, loc) for a, loc in zip(vel, (FACE_X, FACE_Y, FACE_Z))]
    pp = own(p, CELL)
    fu = [own(a, loc) for a, loc in zip(force, (FACE_X, FACE_Y, FACE_Z))] if force is not None else None
    tl = dvel is not None
    du = [own(a, loc) for a, loc in zip(dvel, (FACE_X, FACE_Y, FACE_Z))] if tl else None
    dpp = own(dp, CELL) if tl else None
    dfu = [own(a, loc) for a, loc in zip(dforce, (FACE_X, FACE_Y, FACE_Z))] if dforce is not None else None
    none3 = [None, None, None]
    ke, dke = C.c_double(0), C.c_double(0)
    _call("ofref_tg_run_tl", C.byref(c), C.c_int(scheme), C.c_int(steps),
          *(map(_p, fu) if fu else none3), *map(_p, u), _p(pp), *(map(_p, dfu) if dfu else none3),
          *(map(_p, du) if tl else none3), _p(dpp) if tl else None, C.byref(ke), C.byref(dke))
    return u, pp, ke.value, du, (dke.value if tl else None)


def kinetic_energy(g, vel):
    out = C.c_double(0)
    _call("ofref_kinetic_energy", C.byref(g), *map(_p, vel), C.byref(out))
    return out.value


def cavity_config(nx=128, ny=128, reynolds=100.0, lid_velocity=1.0, cfl_limit=0.5, max_steps=8000,
                  min_steps=200, steady_tolerance=1e-7, max_iterations=200,
                  tolerance=1e-10) -> CavityConfig:
    return CavityConfig(nx, ny, reynolds, lid_velocity, cfl_limit, max_steps, min_steps,
                        steady_tolerance, max_iterations, tolerance)


def cavity_grid(c: CavityConfig) -> Grid:
    return Grid(c.nx, c.ny, 1, 1.0 / c.nx, 1.0 / c.ny, 1.0)


def cavity_init(c: CavityConfig) -> State:
    s = State(cavity_grid(c))
    dt = C.c_double(0)
    _call("ofref_cavity_init", C.byref(c), *map(_p, s.vel), _p(s.p), C.byref(dt))
    s.metrics.dt = dt.value
    return s


def cavity_steps(c: CavityConfig, s: State, steps: int, stop_on_steady=False):
    _call("ofref_cavity_steps", C.byref(c), C.c_int(steps), C.c_int(int(stop_on_steady)),
          *map(_p, s.vel), _p(s.p), *map(_p, s.adv), C.byref(s.has_prev), C.byref(s.metrics))
    return s


def channel_config(case="couette", nx=128, ny=128, viscosity=0.1, top_velocity=1.0, pressure_drop=0.8,
                   cfl_limit=0.5, steps=8, max_iterations=200, tolerance=1e-10) -> ChannelConfig:
    return ChannelConfig(int(case == "poiseuille"), nx, ny, viscosity, top_velocity, pressure_drop, cfl_limit,
                         steps, max_iterations, tolerance)


def channel_grid(c: ChannelConfig) -> Grid:
    return Grid(c.nx, c.ny, 1, 1.0 / c.nx, 1.0 / c.ny, 1.0)


def channel_init(c: ChannelConfig) -> State:
    s = State(channel_grid(c))
    dt = C.c_double(0)
    _call("ofref_channel_init", C.byref(c), *map(_p, s.vel), _p(s.p), C.byref(dt))
    s.metrics.dt = dt.value
    return s


def channel_steps(c: ChannelConfig, s: State, steps: int):
    _call("ofref_channel_steps", C.byref(c), C.c_int(steps), *map(_p, s.vel), _p(s.p), *map(_p, s.adv),
          C.byref(s.has_prev), C.byref(s.metrics))
    return s


def channel_profile(c: ChannelConfig, u):
    """x-averaged streamwise profile (ny values) and its relative L2 error to the analytic one."""
    prof = np.zeros(c.ny)
    err = C.c_double(0)
    _call("ofref_channel_profile", C.byref(c), _p(np.ascontiguousarray(u)), _p(prof), C.byref(err))
    return prof, err.value


def mms(n):
    out = np.zeros(3)
    _call("ofref_mms", C.c_int(n), _p(out))
    return dict(gradient_error=out[0], divergence_error=out[1], laplacian_error=out[2])


# ------------------------------------------------------------------ presets
def tg_bcs(nz) -> BCSet:
    t = [PERIODIC] * 6
    if nz <= 1:
        t[4] = t[5] = SYMMETRY
    return bcset(t)


def cavity_bcs(lid=1.0) -> BCSet:
    t = [NO_SLIP, NO_SLIP, NO_SLIP, PRESCRIBED, SYMMETRY, SYMMETRY]
    vel = [[0, 0, 0]] * 6
    vel = [list(v) for v in vel]
    vel[3][0] = lid
    return bcset(t, vel)
