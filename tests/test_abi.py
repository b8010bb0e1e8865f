"""CPU checks of the C ABI boundary: the built library loads and exports every entry point the
header declares, the Python binding lists the same set, and argument validation that happens
before any device work reports the reference's error class and text."""
import ctypes as C
import os
import re

import pytest

import solver_amd as sa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "of_hip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(of_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "of_solve_pressure_poisson" in names
    assert "of_stepper_run" in names
    assert sorted(sa.EXPORTS) == names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(sa.LIB_PATH):
        pytest.fail(f"{sa.LIB_PATH} not built; run python -m solver_amd.build")
    lib = C.CDLL(sa.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_host_side_validation_messages():
    # no device work: these fail in argument checks before any HIP call
    dt = C.c_double()
    n = C.c_int()
    rc = sa.lib().of_taylor_green_dt(0, 8, 8, C.c_double(0.5), C.c_double(1.0), C.byref(dt), C.byref(n))
    assert rc == sa.OF_INVALID_ARGUMENT
    assert sa.lib().of_last_error().decode() == "Taylor-Green config expects positive nx, ny, and nz"
    dt_val, steps = sa.taylor_green_dt(128, 128, 1, 0.5, 1.0)
    L = 2 * 3.141592653589793
    upper = 0.5 / (128 / L + 128 / L)
    assert steps == int(-(-1.0 // upper))
    assert dt_val == pytest.approx(1.0 / steps, rel=1e-15)
    b = sa.bc_set([sa.PERIODIC, sa.NO_SLIP, 4, 4, 4, 4])
    out = sa.PressureBCSet()
    with pytest.raises(sa.InvalidArgument, match="periodic velocity boundaries must be paired on both sides"):
        sa.call("of_derive_pressure_correction_boundary_conditions", C.byref(b), C.byref(out))
    b = sa.bc_set([sa.NO_SLIP, sa.FIXED_PRESSURE, sa.PERIODIC, sa.PERIODIC, sa.SYMMETRY, sa.PRESCRIBED])
    sa.call("of_derive_pressure_correction_boundary_conditions", C.byref(b), C.byref(out))
    assert [out.faces[f].type for f in range(6)] == [sa.P_NEUMANN, sa.P_DIRICHLET, sa.P_PERIODIC, sa.P_PERIODIC,
                                                     sa.P_NEUMANN, sa.P_NEUMANN]
