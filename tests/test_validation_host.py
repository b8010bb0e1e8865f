"""Host logic of the validation suite (solver_amd/validation/run_validation_suite.py): report
parsing, observed orders, VTK read-back, the gates and the report files, driven by stand-in
tools that print fixed reports (no device work)."""
import math
import os
import stat
import sys

import numpy as np
import pytest

from solver_amd.validation import run_validation_suite as vs

FAKE = r'''#!{python}
import sys, math
name = "{name}"
args = sys.argv[1:]
if name == "solver_operator_mms":
    n = int(args[0])
    e = 1.0 / n ** 2
    print(f"resolution: {{n}}\ngradient_l2_error: {{e}}\ndivergence_l2_error: {{2 * e}}\nlaplacian_l2_error: {{4 * e}}")
elif name == "solver_taylor_green" and "--vtk-out" in args:
    cfg = open(args[0]).read()
    cfl = float(cfg.split("cfl_limit = ")[1].split()[0])
    vtk = args[args.index("--vtk-out") + 1]
    with open(vtk, "w") as f:
        f.write("# vtk DataFile Version 3.0\nx\nASCII\nDATASET STRUCTURED_POINTS\nVECTORS velocity double\n")
        for k in range(128 * 128):
            base = math.sin(k * 0.001) + 2.0
            f.write(f"{{base + cfl ** 2}} {{base}} 0\n")
        f.write("SCALARS pressure double 1\n")
    print(f"dt: {{cfl * 0.02}}\nnormalized_energy_error: 7e-4\nvelocity_relative_l2_error: 2.6e-3\ndivergence_l2: 1e-15")
else:
    status = "{status}"
    print(f"reference_dataset: ds\nprofile_relative_l2_error: 1e-3\nmax_reference_relative_error: 3e-3\n"
          f"normalized_energy_error: 1e-3\ndivergence_l2: 2e-15\nbenchmark_status: {{status}}")
    sys.exit(0 if status == "pass" else 1)
'''


def fake_tools(tmp_path, status="pass"):
    d = tmp_path / "bin"
    d.mkdir()
    for name in ["solver_operator_mms", "solver_taylor_green", "solver_cavity", "solver_channel"]:
        p = d / name
        p.write_text(FAKE.format(python=sys.executable, name=name, status=status))
        p.chmod(p.stat().st_mode | stat.S_IEXEC)
    return d


def test_orders_and_differences():
    assert vs.observed_order(4.0, 1.0) == pytest.approx(2.0)
    a = np.array([1.0, 2.0, 3.0])
    assert vs.relative_rms_difference(a * 1.01, a) == pytest.approx(0.01)
    with pytest.raises(ValueError):
        vs.relative_rms_difference(a, a[:2])
    assert vs.parse_report("a: 1\nnoise\nb: x: y\n") == {"a": "1", "b": "x: y"}


def test_suite_passes_with_second_order_tools(tmp_path):
    out = tmp_path / "out"
    rc = vs.main(["--bin-dir", str(fake_tools(tmp_path)), "--output-dir", str(out)])
    assert rc == 0
    text = (out / "summary.md").read_text()
    assert "Overall status: PASS" in text
    for name in ["operator_spatial", "taylor_green_temporal", "benchmark_summary", "mass_conservation"]:
        assert (out / f"{name}.csv").exists() and (out / f"{name}.md").exists()
    assert (out / "operator_spatial.svg").read_text().startswith("<svg")
    # the stand-in u fields differ by cfl^2: u = base + cfl^2, v = base, w = 0
    base = np.sin(np.arange(128 * 128) * 0.001) + 2.0

    def field(cfl):
        return np.stack([base + cfl ** 2, base, 0 * base], axis=1).ravel()

    d = [vs.relative_rms_difference(field(a), field(b)) for a, b in [(1.0, 0.5), (0.5, 0.25)]]
    rows = (out / "taylor_green_temporal.csv").read_text().splitlines()
    assert math.isclose(float(rows[2].split(",")[-1]), math.log2(d[0] / d[1]), rel_tol=1e-9)
    assert "| 64 | 2.441406e-04 |" in (out / "operator_spatial.md").read_text()


def test_suite_fails_on_a_failed_benchmark(tmp_path):
    out = tmp_path / "out"
    rc = vs.main(["--bin-dir", str(fake_tools(tmp_path, status="fail")), "--output-dir", str(out)])
    assert rc == 1
    text = (out / "summary.md").read_text()
    assert "Overall status: FAIL" in text and "- Benchmark thresholds: fail" in text
