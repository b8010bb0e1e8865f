"""solver_amd: MI355X (gfx950) incompressible projection step behind the C ABI of include/of_hip.h.

This module binds libsolver_amd.so with ctypes. Device memory and streams come from PyTorch
(ROCm build); all arithmetic runs in the library's HIP kernels. There is no CPU fallback:
importing works anywhere, but every call that computes needs the built library and a HIP
device and raises otherwise.

Field tensors are torch.float64 CUDA tensors holding the full ghosted storage of the
reference layout (shape (sz, sy, sx), flat index i + sx*(j + sy*k)), see include/of_hip.h.
"""
from __future__ import annotations

import ctypes as C
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsolver_amd.so")

OF_OK, OF_INVALID_ARGUMENT, OF_RUNTIME_ERROR, OF_HIP_ERROR = 0, -1, -2, -3
CELL, FACE_X, FACE_Y, FACE_Z = 0, 1, 2, 3
NO_SLIP, PRESCRIBED, SYMMETRY, FIXED_PRESSURE, PERIODIC = 0, 1, 2, 3, 4
P_NEUMANN, P_DIRICHLET, P_PERIODIC = 0, 1, 2
PREDICTOR_ADI, PREDICTOR_CG = 0, 1
FP64, MIXED = 0, 1
TAYLOR_GREEN, LID_DRIVEN_CAVITY, CHANNEL_COUETTE, CHANNEL_POISEUILLE = 0, 1, 2, 3


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double)]


class BC(C.Structure):
    _fields_ = [("type", C.c_int), ("velocity", C.c_double * 3), ("pressure", C.c_double),
                ("pressure_gradient", C.c_double)]


class BCSet(C.Structure):
    _fields_ = [("faces", BC * 6)]


class PressureBC(C.Structure):
    _fields_ = [("type", C.c_int), ("value", C.c_double), ("gradient", C.c_double)]


class PressureBCSet(C.Structure):
    _fields_ = [("faces", PressureBC * 6)]


class ProjectionOptions(C.Structure):
    _fields_ = [("dt", C.c_double), ("density", C.c_double), ("poisson_max_iterations", C.c_int),
                ("poisson_tolerance", C.c_double)]


class PoissonDiagnostics(C.Structure):
    _fields_ = [("iterations", C.c_int), ("initial_residual_l2", C.c_double),
                ("final_residual_l2", C.c_double), ("relative_residual", C.c_double),
                ("removed_rhs_mean", C.c_double), ("multigrid_levels", C.c_int),
                ("coarse_unknowns", C.c_int), ("pre_smoothing_steps", C.c_int),
                ("post_smoothing_steps", C.c_int), ("converged", C.c_int),
                ("zero_mean_enforced", C.c_int), ("history_length", C.c_int)]


class ProjectionDiagnostics(C.Structure):
    _fields_ = [("rhs_l2", C.c_double), ("divergence_l2_before", C.c_double),
                ("divergence_l2_after", C.c_double), ("divergence_linf_after", C.c_double),
                ("pressure_mean", C.c_double), ("pressure_solve", PoissonDiagnostics)]


class CflDiagnostics(C.Structure):
    _fields_ = [("dt", C.c_double), ("max_cfl", C.c_double), ("max_u", C.c_double),
                ("max_v", C.c_double), ("max_w", C.c_double)]


class StepperConfig(C.Structure):
    _fields_ = [("flow_case", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("viscosity", C.c_double), ("dt", C.c_double), ("scheme", C.c_int),
                ("poisson_max_iterations", C.c_int), ("poisson_tolerance", C.c_double),
                ("precision", C.c_int), ("lid_velocity", C.c_double), ("pressure_drop", C.c_double)]


class StepMetrics(C.Structure):
    _fields_ = [("step", C.c_int), ("time", C.c_double), ("dt", C.c_double), ("max_cfl", C.c_double),
                ("max_velocity_change", C.c_double), ("divergence_l2", C.c_double),
                ("max_divergence_l2", C.c_double), ("pressure_iterations", C.c_int),
                ("pressure_relative_residual", C.c_double), ("divergence_linf", C.c_double),
                ("max_divergence_linf", C.c_double)]


class SolverError(RuntimeError):
    """Raised for OF_RUNTIME_ERROR / OF_HIP_ERROR (the reference's std::runtime_error)."""


class InvalidArgument(ValueError):
    """Raised for OF_INVALID_ARGUMENT (the reference's std::invalid_argument)."""


_lib = None

# exported C ABI (include/of_hip.h)
EXPORTS = [
    "of_context_create", "of_context_destroy", "of_context_synchronize", "of_last_error", "of_build_info",
    "of_profile_kernel", "of_profile_read", "of_profile_durations", "of_profile_stride",
    "of_active_rms_difference",
    "of_apply_velocity_boundary_conditions", "of_apply_pressure_boundary_conditions",
    "of_apply_total_pressure_boundary_conditions", "of_derive_pressure_correction_boundary_conditions",
    "of_compute_gradient", "of_compute_divergence", "of_compute_laplacian", "of_compute_advection_term",
    "of_compute_diffusion_term", "of_compute_advective_cfl", "of_compute_factorized_correction",
    "of_assemble_predictor_rhs", "of_solve_predictor_adi", "of_solve_predictor_cg", "of_build_pressure_rhs",
    "of_correct_velocity",
    "of_project_velocity", "of_solve_pressure_poisson", "of_solve_pressure_poisson_noalloc",
    "of_build_poisson_rhs_from_pressure",
    "of_apply_poisson_operator", "of_taylor_green_dt", "of_lid_driven_cavity_dt", "of_channel_flow_dt",
    "of_stepper_create",
    "of_stepper_destroy", "of_stepper_initialize", "of_stepper_upload", "of_stepper_download",
    "of_stepper_device_fields", "of_stepper_run", "of_stepper_run_batched", "of_stepper_kinetic_energy",
    "of_stepper_set_forcing", "of_stepper_set_predictor",
    "of_apply_velocity_boundary_conditions_adjoint", "of_apply_pressure_boundary_conditions_adjoint",
    "of_compute_gradient_adjoint", "of_compute_divergence_adjoint", "of_solve_predictor_adi_adjoint",
    "of_compute_advection_tangent", "of_compute_advection_adjoint",
    "of_adjoint_create", "of_adjoint_destroy", "of_adjoint_set_forcing", "of_adjoint_forward",
    "of_adjoint_final_velocity", "of_adjoint_backward", "of_adjoint_tangent",
]


def lib():
    """Loads the HIP library; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m solver_amd.build`")
        # The library and torch must share one HIP runtime: device pointers come from torch's
        # allocator. torch ships its own libamdhip64.so.7; loading torch first makes the
        # library's NEEDED entry bind to that runtime instead of a second copy from /opt/rocm.
        import torch  # noqa: F401

        _lib = C.CDLL(LIB_PATH)
        _lib.of_last_error.restype = C.c_char_p
        _lib.of_build_info.restype = C.c_char_p
    return _lib


def _check(rc):
    if rc == OF_OK:
        return
    msg = lib().of_last_error().decode()
    if rc == OF_INVALID_ARGUMENT:
        raise InvalidArgument(msg)
    raise SolverError(msg)


def call(name, *args):
    _check(getattr(lib(), name)(*args))


def _torch():
    import torch

    return torch


class DevicePointer(C.c_void_p):
    """A device address that keeps its tensor alive for as long as the pointer object lives
    (an argument list built with ptr(tensor_expression) must not free the tensor before the
    call runs)."""


def ptr(t):
    if t is None:
        return C.c_void_p(None)
    torch = _torch()
    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise InvalidArgument("fields must be contiguous float64 device tensors")
    p = DevicePointer(t.data_ptr())
    p.tensor = t
    return p


# ------------------------------------------------------------------ layout helpers
def storage_shape(nx, ny, nz, loc):
    a = [nx, ny, nz]
    if loc != CELL:
        a[loc - 1] += 1
    return (a[2] + 2, a[1] + 2, a[0] + 2)


def grid(nx, ny, nz, dx, dy, dz) -> Grid:
    return Grid(nx, ny, nz, dx, dy, dz)


def bc_set(types, velocity=None, pressure=None) -> BCSet:
    b = BCSet()
    for f in range(6):
        b.faces[f].type = int(types[f])
        for c in range(3):
            b.faces[f].velocity[c] = 0.0 if velocity is None else float(velocity[f][c])
        b.faces[f].pressure = 0.0 if pressure is None else float(pressure[f])
    return b


def pressure_bc_set(types, values=None, gradients=None) -> PressureBCSet:
    b = PressureBCSet()
    for f in range(6):
        b.faces[f].type = int(types[f])
        b.faces[f].value = 0.0 if values is None else float(values[f])
        b.faces[f].gradient = 0.0 if gradients is None else float(gradients[f])
    return b


def taylor_green_bcs(nz) -> BCSet:
    t = [PERIODIC] * 6
    if nz <= 1:
        t[4] = t[5] = SYMMETRY
    return bc_set(t)


def taylor_green_grid(nx, ny, nz) -> Grid:
    L = 2.0 * math.pi
    return Grid(nx, ny, nz, L / nx, L / ny, L / nz if nz > 1 else 1.0)


def taylor_green_dt(nx, ny, nz, cfl_limit, final_time):
    dt, steps = C.c_double(), C.c_int()
    call("of_taylor_green_dt", nx, ny, nz, C.c_double(cfl_limit), C.c_double(final_time), C.byref(dt),
         C.byref(steps))
    return dt.value, steps.value


def lid_driven_cavity_dt(nx, ny, lid_velocity, cfl_limit):
    dt = C.c_double()
    call("of_lid_driven_cavity_dt", nx, ny, C.c_double(lid_velocity), C.c_double(cfl_limit), C.byref(dt))
    return dt.value


def channel_flow_dt(flow_case, nx, viscosity, top_velocity, pressure_drop, cfl_limit):
    dt = C.c_double()
    call("of_channel_flow_dt", flow_case, nx, C.c_double(viscosity), C.c_double(top_velocity),
         C.c_double(pressure_drop), C.c_double(cfl_limit), C.byref(dt))
    return dt.value


# ------------------------------------------------------------------ context
class Context:
    """Device context: reduction buffers and the stream all launches are queued on."""

    def __init__(self, device: int = 0, stream=None):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("solver_amd needs a HIP device")
        self.device = device
        h = C.c_void_p()
        s = C.c_void_p(stream) if stream else C.c_void_p(None)
        call("of_context_create", device, s, C.byref(h))
        self.h = h

    def close(self):
        """Drops this handle's reference to the native context. Steppers and adjoint runs
        created on it hold references of their own in the library, so the native context lives
        until the last of them is closed too, whatever order the handles are released in
        (a garbage collector may finalise a context before the steppers that use it)."""
        if self.h:
            h, self.h = self.h, None
            call("of_context_destroy", h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        call("of_context_synchronize", self.h)

    def profile_kernel(self, name):
        call("of_profile_kernel", self.h, name.encode() if name else None)

    def profile_read(self):
        n, tot, mn, mx = C.c_int(), C.c_double(), C.c_double(), C.c_double()
        call("of_profile_read", self.h, C.byref(n), C.byref(tot), C.byref(mn), C.byref(mx))
        return dict(launches=n.value, total_ms=tot.value, min_ms=mn.value, max_ms=mx.value)

    def profile_stride(self, stride):
        """Time only every stride-th launch of the profiled kernel."""
        call("of_profile_stride", self.h, int(stride))

    def profile_durations(self, capacity=1 << 16):
        """Per-launch durations (ms) of the profiled kernel since the last read, in launch order."""
        buf = (C.c_double * capacity)()
        n = C.c_int()
        call("of_profile_durations", self.h, buf, capacity, C.byref(n))
        return [buf[i] for i in range(min(n.value, capacity))]

    def run(self, name, *args):
        """Calls one entry point after torch's pending work and waits for it to finish."""
        _torch().cuda.synchronize()
        call(name, self.h, *args)
        self.sync()


class Stepper:
    """Device-resident Taylor-Green / lid-driven-cavity / channel state and step loop.

    lid_velocity is the cavity's lid speed and the Couette channel's top-wall speed;
    pressure_drop is the Poiseuille channel's inlet pressure."""

    def __init__(self, ctx: Context, flow_case, nx, ny, nz, viscosity, dt, poisson_max_iterations=200,
                 poisson_tolerance=1e-10, precision=FP64, scheme=0, lid_velocity=1.0, pressure_drop=0.0):
        self.ctx = ctx
        cfg = StepperConfig(flow_case, nx, ny, nz, viscosity, dt, scheme, poisson_max_iterations,
                            poisson_tolerance, precision, lid_velocity, pressure_drop)
        self.cfg = cfg
        h = C.c_void_p()
        call("of_stepper_create", ctx.h, C.byref(cfg), C.byref(h))
        self.h = h
        self.nx, self.ny, self.nz = nx, ny, nz

    def close(self):
        if self.h:
            h, self.h = self.h, None
            call("of_stepper_destroy", h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initialize(self, convention=0):
        call("of_stepper_initialize", self.h, convention)

    def upload(self, u, v, w, p, au=None, av=None, aw=None, has_prev=0, metrics=None):
        import numpy as np

        def hp(a):
            if a is None:
                return C.c_void_p(None)
            a = np.ascontiguousarray(a, dtype=np.float64)
            self._keep.append(a)
            return a.ctypes.data_as(C.c_void_p)

        self._keep = []
        call("of_stepper_upload", self.h, hp(u), hp(v), hp(w), hp(p), hp(au), hp(av), hp(aw), int(has_prev),
             C.byref(metrics) if metrics is not None else None)
        self._keep = []

    def download(self):
        import numpy as np

        shp = [storage_shape(self.nx, self.ny, self.nz, loc) for loc in (FACE_X, FACE_Y, FACE_Z)]
        vel = [np.zeros(s) for s in shp]
        adv = [np.zeros(s) for s in shp]
        p = np.zeros(storage_shape(self.nx, self.ny, self.nz, CELL))
        hpv = C.c_int()
        m = StepMetrics()
        ps = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        call("of_stepper_download", self.h, *map(ps, vel), ps(p), *map(ps, adv), C.byref(hpv), C.byref(m))
        return vel, p, adv, hpv.value, m

    def run(self, steps) -> StepMetrics:
        m = StepMetrics()
        call("of_stepper_run", self.h, int(steps), C.byref(m))
        return m

    def kinetic_energy(self):
        e = C.c_double()
        call("of_stepper_kinetic_energy", self.h, C.byref(e))
        return e.value

    def device_fields(self):
        arr = (C.c_void_p * 7)()
        call("of_stepper_device_fields", self.h, arr)
        return [arr[i] for i in range(7)]

    def set_forcing(self, f=None):
        """Body force (three float64 device tensors in the velocity layout, or None)."""
        _torch().cuda.synchronize()
        fs = [None] * 3 if f is None else list(f)
        call("of_stepper_set_forcing", self.h, *map(ptr, fs))
        self.ctx.sync()

    def set_predictor(self, kind=PREDICTOR_ADI, tolerance=1e-12, max_iterations=80):
        """Helmholtz predictor of the following steps: PREDICTOR_ADI (exact line solves) or
        PREDICTOR_CG (conjugate gradients on the factorised operator; periodic grids only)."""
        call("of_stepper_set_predictor", self.h, int(kind), C.c_double(tolerance), int(max_iterations))


class Adjoint:
    """Checkpointed discrete adjoint and tangent-linear model of the periodic 3D Taylor-Green
    step (device tensors in the reference storage layout)."""

    def __init__(self, ctx: Context, nx, ny, nz, viscosity, dt, checkpoint_every, poisson_max_iterations=400,
                 poisson_tolerance=1e-10, precision=FP64, scheme=0):
        self.ctx = ctx
        self.cfg = StepperConfig(TAYLOR_GREEN, nx, ny, nz, viscosity, dt, scheme, poisson_max_iterations,
                                 poisson_tolerance, precision, 0.0, 0.0)
        h = C.c_void_p()
        call("of_adjoint_create", ctx.h, C.byref(self.cfg), int(checkpoint_every), C.byref(h))
        self.h = h
        self.shape = (nx, ny, nz)

    def close(self):
        if self.h:
            h, self.h = self.h, None
            call("of_adjoint_destroy", h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _new(self, loc):
        torch = _torch()
        return torch.zeros(storage_shape(*self.shape, loc), dtype=torch.float64, device=f"cuda:{self.ctx.device}")

    def set_forcing(self, f=None):
        _torch().cuda.synchronize()
        fs = [None] * 3 if f is None else list(f)
        call("of_adjoint_set_forcing", self.h, *map(ptr, fs))
        self.ctx.sync()

    def forward(self, vel, p=None, steps=1):
        """Runs `steps` steps from vel = (u, v, w) and p; returns (final kinetic energy, metrics)."""
        _torch().cuda.synchronize()
        ke = C.c_double()
        m = StepMetrics()
        call("of_adjoint_forward", self.h, *map(ptr, vel), ptr(p), int(steps), C.byref(ke), C.byref(m))
        return ke.value, m

    def final_velocity(self):
        out = [self._new(loc) for loc in (FACE_X, FACE_Y, FACE_Z)]
        call("of_adjoint_final_velocity", self.h, *map(ptr, out))
        return out

    def backward(self, seed=None, want_pressure=True, want_forcing=True):
        """Gradients w.r.t. (u0, v0, w0), p0 and the forcing field of <seed, u^N> (seed None: of
        the final kinetic energy)."""
        torch = _torch()
        torch.cuda.synchronize()
        gu = [self._new(loc) for loc in (FACE_X, FACE_Y, FACE_Z)]
        gp = self._new(CELL) if want_pressure else None
        gf = [self._new(loc) for loc in (FACE_X, FACE_Y, FACE_Z)] if want_forcing else [None] * 3
        sd = [None] * 3 if seed is None else [s.contiguous() for s in seed]
        call("of_adjoint_backward", self.h, *map(ptr, sd), *map(ptr, gu), ptr(gp), *map(ptr, gf))
        return gu, gp, (gf if want_forcing else None)

    def tangent(self, dvel=None, dp=None, dforce=None):
        """Tangent-linear run from the last forward pass's initial state: returns the final velocity
        perturbation and the directional derivative of the final kinetic energy."""
        _torch().cuda.synchronize()
        dv = [None] * 3 if dvel is None else list(dvel)
        df = [None] * 3 if dforce is None else list(dforce)
        out = [self._new(loc) for loc in (FACE_X, FACE_Y, FACE_Z)]
        dke = C.c_double()
        call("of_adjoint_tangent", self.h, *map(ptr, dv), ptr(dp), *map(ptr, df), *map(ptr, out), C.byref(dke))
        return out, dke.value
