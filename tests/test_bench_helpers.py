"""bench.py's reporting helpers (CPU): the placement summary beside `value`,
the statement of what `value` is, and the candidate budget."""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_placement_summary_first_median_kept(bench):
    p = {"method": "window", "slack_gib": 4, "step_gib": 2, "glups_each": [12.0, 14.0, 13.0], "kept": 1}
    s = bench.placement_summary(p, 432.0)
    assert (s["glups_first_alloc"], s["glups_median"], s["glups_kept"]) == (12.0, 13.0, 14.0)
    assert s["frac_first_alloc"] == pytest.approx(12.0 * 432 / 8000)
    assert s["frac_kept"] == pytest.approx(14.0 * 432 / 8000)
    # candidates that could not be allocated (None) are left out of the median
    p = {"method": "separate allocations", "trials": 4, "glups_each": [10.0, 11.0, None, None], "kept": 1}
    s = bench.placement_summary(p, 216.0)
    assert s["glups_median"] == pytest.approx(10.5) and s["glups_kept"] == 11.0
    assert bench.placement_summary(None, 432.0) == {}


def test_value_basis_names_the_search(bench):
    assert "first lattice allocation" in bench.value_basis(None)
    w = {"method": "window", "slack_gib": 64, "step_gib": 2, "glups_each": [1.0] * 33, "kept": 0,
         "held_gb_during_search": 97.7, "held_gb_after": 102.0}
    text = bench.value_basis(w)
    assert "33 offsets" in text and "2 GiB apart" in text and "64 GiB" in text and "102.0 GB after" in text
    a = {"method": "separate allocations", "trials": 4, "glups_each": [1.0] * 4, "kept": 2,
         "held_gb_during_search": 116.0, "held_gb_after": 29.0}
    assert "fastest of 4 lattice allocations" in bench.value_basis(a)


def test_trials_for_budget(bench):
    # 512^3 AA: FP64 29 GB, FP32 14.5 GB, FP16 7.2 GB per lattice
    assert bench.trials_for_budget(512, "f64", "aa", 8, 128.0)[0] == 4
    assert bench.trials_for_budget(512, "f32", "aa", 8, 128.0)[0] == 8
    assert bench.trials_for_budget(512, "f16", "aa", 12, 128.0)[0] == 12
    assert bench.trials_for_budget(512, "f64", "pull", 8, 128.0)[0] == 2
    assert bench.trials_for_budget(512, "f64", "aa", 8, 1.0)[0] == 1
