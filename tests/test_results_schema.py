"""Results-JSON schema of benchmarks/benchmark_gray_scott.py:458-475 and the
benchmark's inputs (no GPU).  The fixture tests/golden/
reference_gray_scott_results.json is the reference's own results file
(benchmarks/results/gray_scott.json), kept as data."""
import copy
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(__file__)


def _ref():
    with open(os.path.join(HERE, "golden", "reference_gray_scott_results.json")) as f:
        return json.load(f)


def test_reference_results_validate_and_roundtrip(tmp_path):
    from moljax_amd.gray_scott_benchmark import compare, load_results, save_results, validate_results
    ref = _ref()
    validate_results(ref)
    p = save_results(ref, tmp_path / "r" / "gray_scott.json")
    assert load_results(p) == ref
    ours = copy.deepcopy(ref)
    ours["methods"] = [dict(m, time_s=m["time_s"] / 4) for m in ref["methods"][2:]]
    sp = compare(ours, ref)
    assert sp == {"IMEX-Strang": pytest.approx(4.0), "ETDRK4": pytest.approx(4.0)}


@pytest.mark.parametrize("breakage", ["extra_top", "missing_method_key", "config_type", "bool_as_int"])
def test_schema_rejects(breakage):
    from moljax_amd.gray_scott_benchmark import validate_results
    bad = _ref()
    if breakage == "extra_top":
        bad["extra"] = 1
    elif breakage == "missing_method_key":
        del bad["methods"][0]["iqr_s"]
    elif breakage == "config_type":
        bad["config"]["N"] = "256"
    else:
        bad["config"]["n_reps"] = True
    with pytest.raises(ValueError):
        validate_results(bad)


def test_benchmark_inputs():
    """benchmark_gray_scott.py:45-70: the seed-42 draws, the perturbed
    square and the -k^2 symbol (its largest magnitude (N/2 2 pi)^2 for two
    axes is the 1.29e6 the script quotes at :87)."""
    from moljax_amd.gray_scott_benchmark import benchmark_ic, spectral_symbol
    u0, v0 = benchmark_ic(256)
    assert u0.shape == v0.shape == (256, 256)
    assert np.all(u0[118:138, 118:138] == 0.5) and np.all(v0[118:138, 118:138] == 0.25)
    mask = np.ones((256, 256), bool)
    mask[118:138, 118:138] = False
    assert 0.95 <= u0[mask].min() and u0[mask].max() <= 1.0
    assert 0.25 <= v0[mask].min() and v0[mask].max() <= 0.30
    np.random.seed(42)
    first = np.random.rand()
    assert u0[0, 0] == 1.0 - 0.5 * first * 0.1
    lam = spectral_symbol(256)
    assert lam[0] == 0.0 and lam.min() == -(128 * 2 * np.pi) ** 2
    assert abs(2 * (128 * 2 * np.pi) ** 2 - 1.29e6) / 1.29e6 < 0.01
