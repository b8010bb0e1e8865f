"""The C++ interface builds, and its host-only parts behave as the reference's: layouts,
boundary descriptors, the .cfg surface with its error messages, the cavity envelope check
(solver_amd/cpp/tests/host_checks.cpp). No device work."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "solver_amd", "cpp")
BIN = os.path.join(CPP, "build", "bin")
TOOLS = ["solver_taylor_green", "solver_cavity", "solver_channel", "solver_pressure_profile", "solver_advection_profile",
         "solver_operator_mms", "api_roundtrip", "host_checks", "io_check"]


def _build():
    subprocess.run(["make", "-C", CPP, "-j8"], check=True, capture_output=True, timeout=600)


def test_cpp_layer_builds_every_tool():
    _build()
    for t in TOOLS:
        assert os.access(os.path.join(BIN, t), os.X_OK), t


def test_host_checks():
    _build()
    r = subprocess.run([os.path.join(BIN, "host_checks")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all host checks passed" in r.stdout


def test_tool_usage_errors_name_the_tool():
    _build()
    r = subprocess.run([os.path.join(BIN, "solver_taylor_green"), "--bogus"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1
    assert "solver_taylor_green failed: unsupported argument: --bogus" in r.stderr
    r = subprocess.run([os.path.join(BIN, "solver_cavity"), "/nonexistent.cfg"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1
    assert "solver_cavity failed: unable to open cavity config: /nonexistent.cfg" in r.stderr
