"""The Gray-Scott benchmark's own schemes on the GPU vs the float64 oracle:
its splitting (benchmark_gray_scott.py:268-289, mjx_gs_imex scheme 3) and its
ETDRK4 with f2 as written (:336-417, mjx_etdrk4_coeffs_form 1), from its
seed-42 initial condition on 256^2.  Tolerances (relative L2): float64
1e-12, float32 1e-5.  The results dict follows the script's schema."""
import math

import numpy as np
import pytest
import torch

from oracle import etdrk4 as oe
from oracle import gray_scott as ogs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["IMEX-Strang", "ETDRK4"])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 1e-5)])
def test_benchmark_methods_match_oracle(gpu, name, dtype, tol):
    from moljax_amd.gray_scott_benchmark import DU, DV, F, K, GrayScottMethod, benchmark_ic, spectral_symbol
    u0, v0 = benchmark_ic(256)
    lam = spectral_symbol(256)
    m = GrayScottMethod(name, 256, dtype)
    m.load(torch.from_numpy(u0).to("cuda", dtype), torch.from_numpy(v0).to("cuda", dtype))
    steps = 20 if name == "IMEX-Strang" else 10
    m.advance(steps)
    if name == "IMEX-Strang":
        ru, rv = oe.benchmark_strang(u0, v0, lam, lam, DU, DV, F, K, 0.5, steps)
    else:
        ru, rv = oe.gs_etdrk4(u0, v0, lam, lam, DU, DV, F, K, 1.0, steps, cox_matthews=False)
    eu = ogs.rel_l2(m.u.double().cpu().numpy(), ru)
    ev = ogs.rel_l2(m.v.double().cpu().numpy(), rv)
    print(f"{name} {dtype} {steps} steps: rel_l2 u={eu:.2e} v={ev:.2e}", flush=True)
    assert eu <= tol and ev <= tol


def test_benchmark_strang_chunks_and_batch_rows(gpu):
    """Splitting the run over calls (the closing / opening half-steps then
    meet at a call boundary instead of inside one row pass) gives the same
    trajectory to rounding."""
    from moljax_amd.gray_scott_benchmark import GrayScottMethod, benchmark_ic
    u0, v0 = benchmark_ic(256)
    a = GrayScottMethod("IMEX-Strang", 256, torch.float64)
    b = GrayScottMethod("IMEX-Strang", 256, torch.float64)
    for m in (a, b):
        m.load(torch.from_numpy(u0).cuda(), torch.from_numpy(v0).cuda())
    a.advance(12)
    for n in (5, 1, 6):
        b.advance(n)
    assert ogs.rel_l2(b.u.cpu().numpy(), a.u.cpu().numpy()) < 1e-14
    assert ogs.rel_l2(b.v.cpu().numpy(), a.v.cpu().numpy()) < 1e-14


def test_run_benchmark_schema(gpu, tmp_path):
    from moljax_amd.gray_scott_benchmark import load_results, run_benchmark, save_results
    res = run_benchmark(t_final=5.25, n_reps=2, dtype=torch.float32, warmup_steps=2)
    p = save_results(res, tmp_path / "gray_scott.json")
    back = load_results(p)
    names = [m["name"] for m in back["methods"]]
    assert names == ["IMEX-Strang", "ETDRK4"]
    for m in back["methods"]:
        assert m["steps"] == math.ceil(5.25 / m["dt"]) and m["t_final"] == m["steps"] * m["dt"] >= 5.25
        assert m["time_s"] > 0 and m["iqr_s"] >= 0
    assert back["config"]["dtype"] == "float32" and back["config"]["N"] == 256
    with pytest.raises(ValueError):
        run_benchmark(t_final=1.0, n_reps=1, methods=("RK4 (explicit)",))
