"""The Gray-Scott benchmark's own schemes on the GPU vs the float64 oracle:
its splitting (benchmark_gray_scott.py:268-289, mjx_gs_imex scheme 3), its
ETDRK4 with f2 as written (:336-417, mjx_etdrk4_coeffs_form 1), its
pseudo-spectral RK4 (:89-127, mjx_gs_rk4_spectral) and its preconditioned
Crank-Nicolson iteration (:162-229, mjx_gs_cn_fixed_point, iteration counts
identical), from its seed-42 initial condition on 256^2.  Tolerances
(relative L2): float64 1e-12, float32 1e-5.  The results dict follows the
script's schema."""
import math

import numpy as np
import pytest
import torch

from oracle import etdrk4 as oe
from oracle import gray_scott as ogs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("spg", [0, 4])
@pytest.mark.parametrize("name", ["IMEX-Strang", "ETDRK4"])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 1e-5)])
def test_benchmark_methods_match_oracle(gpu, name, dtype, tol, spg):
    """Direct launches (spg 0) and hipGraph replays of 4 steps (spg 4)."""
    from moljax_amd.gray_scott_benchmark import DU, DV, F, K, GrayScottMethod, benchmark_ic, spectral_symbol
    u0, v0 = benchmark_ic(256)
    lam = spectral_symbol(256)
    m = GrayScottMethod(name, 256, dtype, steps_per_graph=spg)
    m.load(torch.from_numpy(u0).to("cuda", dtype), torch.from_numpy(v0).to("cuda", dtype))
    steps = 20 if name == "IMEX-Strang" else 10
    m.advance(steps)
    if name == "IMEX-Strang":
        ru, rv = oe.benchmark_strang(u0, v0, lam, lam, DU, DV, F, K, 0.5, steps)
    else:
        ru, rv = oe.gs_etdrk4(u0, v0, lam, lam, DU, DV, F, K, 1.0, steps, cox_matthews=False)
    eu = ogs.rel_l2(m.u.double().cpu().numpy(), ru)
    ev = ogs.rel_l2(m.v.double().cpu().numpy(), rv)
    print(f"{name} {dtype} {steps} steps (graph {spg}): rel_l2 u={eu:.2e} v={ev:.2e}", flush=True)
    assert eu <= tol and ev <= tol


def test_benchmark_strang_chunks_and_batch_rows(gpu):
    """Splitting the run over calls (the closing / opening half-steps then
    meet at a call boundary instead of inside one row pass) gives the same
    trajectory to rounding; so does a batch of two benchmark-splitting
    instances in one graph plan, each bitwise equal to its lone plan."""
    from moljax_amd.gray_scott_benchmark import GrayScottMethod, benchmark_ic
    u0, v0 = benchmark_ic(256)
    a = GrayScottMethod("IMEX-Strang", 256, torch.float64, steps_per_graph=0)
    b = GrayScottMethod("IMEX-Strang", 256, torch.float64, steps_per_graph=0)
    for m in (a, b):
        m.load(torch.from_numpy(u0).cuda(), torch.from_numpy(v0).cuda())
    a.advance(12)
    for n in (5, 1, 6):
        b.advance(n)
    assert ogs.rel_l2(b.u.cpu().numpy(), a.u.cpu().numpy()) < 1e-14
    assert ogs.rel_l2(b.v.cpu().numpy(), a.v.cpu().numpy()) < 1e-14
    from moljax_amd.batch import GrayScottBatch
    from moljax_amd.gray_scott_benchmark import DU, DV
    mix = GrayScottBatch([64, 128, 64], [0.03, 0.035, 0.04], [0.06, 0.065, 0.062], DU, DV, 0.5,
                         scheme="benchmark_strang", dtype=torch.float64, steps_per_graph=3)
    rng = np.random.default_rng(1)
    ics = [(1 - 0.1 * rng.random((n, n)), 0.25 + 0.1 * rng.random((n, n))) for n in (64, 128, 64)]
    for i, (uu, vv) in enumerate(ics):
        mix.set_state(i, torch.from_numpy(uu), torch.from_numpy(vv))
    mix.run(8)
    lam128 = oe.spectral_symbol_1d(128)
    ru, _ = oe.benchmark_strang(ics[1][0], ics[1][1], lam128, lam128, DU, DV, 0.035, 0.065, 0.5, 8)
    assert ogs.rel_l2(mix.state(1)[0].cpu().numpy(), ru) < 1e-12
    alone = GrayScottBatch([64], [0.04], [0.062], DU, DV, 0.5, scheme="benchmark_strang", dtype=torch.float64,
                           steps_per_graph=3)
    alone.set_state(0, torch.from_numpy(ics[2][0]), torch.from_numpy(ics[2][1]))
    alone.run(8)
    assert torch.equal(alone.state(0)[0], mix.state(2)[0]) and torch.equal(alone.state(0)[1], mix.state(2)[1])
    mix.close()
    alone.close()


@pytest.mark.parametrize("spg", [0, 3])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 1e-5)])
def test_rk4_and_cn_match_oracle(gpu, dtype, tol, spg):
    """RK4 (20 steps at the script's dt) and CN (3 steps of dt 0.5); in FP64
    the CN iteration count of every step equals the oracle's.  spg 3 replays
    captured 3-step sequences (twice, then one step directly)."""
    from moljax_amd.gray_scott_benchmark import (DU, DV, F, K, GrayScottMethod, benchmark_ic, method_dt,
                                                 spectral_symbol)
    u0, v0 = benchmark_ic(256)
    lam = spectral_symbol(256)
    rk = GrayScottMethod("RK4 (explicit)", 256, dtype, steps_per_graph=spg)
    rk.load(torch.from_numpy(u0).to("cuda", dtype), torch.from_numpy(v0).to("cuda", dtype))
    rk.advance(7)
    ru, rv = u0, v0
    for _ in range(7):
        ru, rv = oe.benchmark_rk4_step(ru, rv, lam, lam, DU, DV, F, K, method_dt("RK4 (explicit)"))
    eu = ogs.rel_l2(rk.u.double().cpu().numpy(), ru)
    ev = ogs.rel_l2(rk.v.double().cpu().numpy(), rv)
    print(f"RK4 {dtype} graph {spg}: rel_l2 u={eu:.2e} v={ev:.2e}", flush=True)
    assert eu <= tol and ev <= tol
    cn = GrayScottMethod("CN (Newton-Krylov)", 256, dtype, steps_per_graph=spg)
    cn.load(torch.from_numpy(u0).to("cuda", dtype), torch.from_numpy(v0).to("cuda", dtype))
    cn.advance(7)
    cu, cv, its = u0, v0, []
    for _ in range(7):
        cu, cv, it = oe.benchmark_cn_step(cu, cv, lam, lam, DU, DV, F, K, 0.5)
        its.append(it)
    eu = ogs.rel_l2(cn.u.double().cpu().numpy(), cu)
    ev = ogs.rel_l2(cn.v.double().cpu().numpy(), cv)
    got = int(cn.iters.item())
    print(f"CN {dtype} graph {spg}: rel_l2 u={eu:.2e} v={ev:.2e} iterations {got} (oracle {sum(its)}: {its})",
          flush=True)
    assert eu <= tol and ev <= tol
    if dtype == torch.float64:
        assert got == sum(its)
    else:
        assert sum(its) <= got <= 10 * 7


def test_run_benchmark_schema(gpu, tmp_path):
    from moljax_amd.gray_scott_benchmark import METHODS, load_results, run_benchmark, save_results
    res = run_benchmark(t_final=5.25, n_reps=2, dtype=torch.float32, warmup_steps=2, steps_per_graph=3)
    p = save_results(res, tmp_path / "gray_scott.json")
    back = load_results(p)
    names = [m["name"] for m in back["methods"]]
    assert names == list(METHODS)
    for m in back["methods"]:
        assert m["steps"] == math.ceil(5.25 / m["dt"]) and m["t_final"] == m["steps"] * m["dt"] >= 5.25
        assert m["time_s"] > 0 and m["iqr_s"] >= 0
    assert back["config"]["dtype"] == "float32" and back["config"]["N"] == 256
    with pytest.raises(ValueError):
        run_benchmark(t_final=1.0, n_reps=1, methods=("Heun",))
