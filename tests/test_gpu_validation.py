"""The validation suite (solver_amd/validation/run_validation_suite.py) on the device: every gate
passes, and the figures agree with the reference's own published run of its suite
(validation/latest/*.csv in the reference checkout, fixture validation_latest). The reference
printed its figures with 6 significant digits, which sets the comparison tolerances."""
import csv

import pytest

from solver_amd.validation import run_validation_suite as vs
from tests.golden import fixtures as fx

pytestmark = pytest.mark.gpu


def rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def close(a, b, rel):
    a, b = float(a), float(b)
    return abs(a - b) <= rel * abs(b)


def test_validation_suite_passes_and_matches_the_reference(tmp_path):
    out = tmp_path / "validation"
    print("running the validation suite", flush=True)
    assert vs.main(["--output-dir", str(out)]) == 0
    assert "Overall status: PASS" in (out / "summary.md").read_text()
    ref = fx.load("validation_latest").meta["tables"]
    for mine, theirs in zip(rows(out / "operator_spatial.csv"), ref["operator_spatial"]):
        assert int(mine["resolution"]) == int(float(theirs["resolution"]))
        for k in ["gradient_l2_error", "divergence_l2_error", "laplacian_l2_error"]:
            assert close(mine[k], theirs[k], 2e-6), k
    for mine, theirs in zip(rows(out / "taylor_green_temporal.csv"), ref["taylor_green_temporal"]):
        for k in ["dt", "normalized_energy_error", "velocity_relative_l2_error"]:
            assert close(mine[k], theirs[k], 2e-6), k
        assert float(mine["divergence_l2"]) <= 1e-10
        if theirs["self_difference_to_finer"]:
            # both sides difference fields printed with 6 significant digits
            assert close(mine["self_difference_to_finer"], theirs["self_difference_to_finer"], 1e-3)
        if theirs["self_convergence_order"]:
            assert abs(float(mine["self_convergence_order"]) - float(theirs["self_convergence_order"])) <= 2e-3
    for mine, theirs in zip(rows(out / "benchmark_summary.csv"), ref["benchmark_summary"]):
        assert mine["case"] == theirs["case"] and mine["reference_dataset"] == theirs["reference_dataset"]
        assert mine["benchmark_status"] == theirs["benchmark_status"] == "pass"
        assert close(mine["metric_value"], theirs["metric_value"], 2e-5), mine["case"]
        assert float(mine["divergence_l2"]) <= 1e-10
