"""IMEX-Euler / SSPRK2 steps and the adaptive IMEX driver on the GPU vs the
float64 oracle.

Adaptive parity: with a float64 state the device controller reproduces the
oracle's accept/reject sequence decision for decision (same n_accepted,
n_rejected, status, write index) and its dt / t histories to 1e-10
relative; the saved states agree to 1e-10.  A batch of instances equals the
same instances run alone, bitwise.  A float32 state with the float64 norm
is compared on the decision sequence (first divergence printed) and on the
final time and status.
"""
import numpy as np
import pytest
import torch

from oracle import adaptive as oad
from oracle import gray_scott as ogs
from oracle import spectral as osp

pytestmark = pytest.mark.gpu

L = 2.5
GS = dict(Du=0.16, Dv=0.08, F=0.04, k=0.06)  # create_gray_scott_periodic_fft defaults
PID = dict(atol=1e-4, rtol=1e-3, dt_min=1e-6, dt_max=5.0)  # examples/gray_scott_2d.py:165-168


def _np(t):
    return t.detach().double().cpu().numpy()


def _model(n, dtype, **kw):
    from moljax_amd.core import Grid2D, create_gray_scott_periodic_fft
    grid = Grid2D.uniform(n, n, 0.0, L, 0.0, L)
    p = dict(GS, **kw)
    return create_gray_scott_periodic_fft(grid, dtype=dtype, **p), p


def _lam(n):
    return osp.laplacian_symbol_2d(n, n, L / n, L / n)


def _dev(y0, dtype):
    return {k: torch.from_numpy(v).to("cuda", dtype) for k, v in y0.items()}


@pytest.mark.parametrize("scheme", ["euler", "ssprk2"])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 2e-6)])
def test_imex_euler_ssprk2_steps(gpu, scheme, dtype, tol):
    from moljax_amd.core.stepping import imex_euler_step, imex_ssprk2_step
    n, dt = 64, 0.5
    (model, cache, diff), p = _model(n, dtype)
    y0 = ogs.initial_condition(n, n, seed=2, square=16)
    ostep = ogs.imex_euler_step if scheme == "euler" else ogs.imex_ssprk2_step
    gstep = imex_euler_step if scheme == "euler" else imex_ssprk2_step
    ref, y = y0, _dev(y0, dtype)
    for _ in range(3):
        ref = ostep(ref, dt, p["F"], p["k"], p["Du"], p["Dv"], _lam(n))
        y = gstep(model, y, 0.0, dt, cache, diff)
    for k in ("u", "v"):
        assert ogs.rel_l2(_np(y[k]), ref[k]) <= tol, k
        assert np.max(np.abs(_np(y[k]) - ref[k])) <= 10 * tol * np.max(np.abs(ref[k]))  # ghosts included


def test_fixed_dt_euler_driver(gpu):
    from moljax_amd.core.stepping import integrate_imex_fixed_dt
    n, dt = 64, 0.25
    (model, cache, diff), p = _model(n, torch.float64)
    y0 = ogs.initial_condition(n, n, seed=4, square=16)
    _, _, fin = integrate_imex_fixed_dt(model, _dev(y0, torch.float64), 0.0, 2.0, dt, cache, diff,
                                        use_strang=False, save_every=3)
    ref = y0
    for _ in range(ogs.fixed_dt_n_steps(0.0, 2.0, dt)):
        ref = ogs.imex_euler_step(ref, dt, p["F"], p["k"], p["Du"], p["Dv"], _lam(n))
    assert ogs.rel_l2(_np(fin["u"]), ref["u"]) < 1e-12 and ogs.rel_l2(_np(fin["v"]), ref["v"]) < 1e-12


def _run_both(n, seed, t_end, dt0, use_strang=True, dtype=torch.float64, max_steps=400, save_every=1, **pidkw):
    from moljax_amd.core.dt_policy import PIDParams
    from moljax_amd.core.stepping import adaptive_integrate_imex
    (model, cache, diff), p = _model(n, dtype)
    y0 = ogs.initial_condition(n, n, seed=seed, square=n // 4)
    pk = dict(PID, **pidkw)
    ref = oad.adaptive_integrate_imex(y0, 0.0, t_end, dt0, p["F"], p["k"], p["Du"], p["Dv"], _lam(n), L / n,
                                      use_strang=use_strang, max_steps=max_steps, save_every=save_every,
                                      pid=oad.PIDParams(**pk))
    print(f"n={n} seed={seed} strang={use_strang} {dtype}: oracle acc={ref.n_accepted} rej={ref.n_rejected} "
          f"status={ref.status}", flush=True)
    res = adaptive_integrate_imex(model, _dev(y0, dtype), 0.0, t_end, dt0, cache, diff, use_strang=use_strang,
                                  max_steps=max_steps, pid_params=PIDParams(**pk), save_every=save_every)
    return ref, res, y0


@pytest.mark.parametrize("use_strang", [True, False])
@pytest.mark.parametrize("n", [64, 48])
def test_adaptive_fp64_matches_oracle(gpu, use_strang, n):
    ref, res, _ = _run_both(n, 1, 20.0, 0.5, use_strang=use_strang)
    assert ref.n_rejected > 0  # the run exercises the reject branch
    assert (res.n_accepted, res.n_rejected, res.status, res.n_steps) == \
        (ref.n_accepted, ref.n_rejected, ref.status, ref.n_steps)
    assert res.t_final == pytest.approx(ref.t_final, rel=1e-12)
    w = ref.n_steps
    assert np.allclose(_np(res.dt_history)[:w], ref.dt_history[:w], rtol=1e-10, atol=0)
    assert np.allclose(_np(res.t_history)[:w], ref.t_history[:w], rtol=1e-10, atol=0)
    for i in (0, 1, w // 2, w - 1):
        for f in ("u", "v"):
            assert ogs.rel_l2(_np(res.y_history[f][i]), ref.y_saved[i][f]) < 1e-10, (i, f)
    for f in ("u", "v"):
        assert ogs.rel_l2(_np(res.y_final[f]), ref.y_final[f]) < 1e-10  # ghosts included


def test_adaptive_save_every(gpu):
    ref, res, _ = _run_both(32, 3, 5.0, 0.2, save_every=4)
    assert (res.n_accepted, res.n_rejected, res.n_steps) == (ref.n_accepted, ref.n_rejected, ref.n_steps)
    w = ref.n_steps
    assert np.allclose(_np(res.t_history)[:w], ref.t_history[:w], rtol=1e-10, atol=0)
    assert ogs.rel_l2(_np(res.y_history["u"][w - 1]), ref.y_saved[w - 1]["u"]) < 1e-10


def test_adaptive_batch_equals_alone(gpu):
    """Three instances, each with its own dt sequence, in one call: each
    equals its single-instance run bitwise."""
    from moljax_amd.core.dt_policy import PIDParams
    from moljax_amd.core.stepping import adaptive_integrate_imex
    n = 64
    (model, cache, diff), p = _model(n, torch.float64)
    ys = [ogs.initial_condition(n, n, seed=s, square=8 + 8 * s) for s in range(3)]
    pid = PIDParams(**PID)
    alone = [adaptive_integrate_imex(model, _dev(y, torch.float64), 0.0, 10.0, 0.5, cache, diff, max_steps=300,
                                     pid_params=pid) for y in ys]
    yb = {f: torch.stack([torch.from_numpy(y[f]) for y in ys]).cuda() for f in ("u", "v")}
    batch = adaptive_integrate_imex(model, yb, 0.0, 10.0, 0.5, cache, diff, max_steps=300, pid_params=pid)
    counts = [(a.n_accepted, a.n_rejected) for a in alone]
    print("per-instance (accepted, rejected):", counts, flush=True)
    assert len(set(counts)) > 1  # the instances really follow different sequences
    for b, a in enumerate(alone):
        assert int(batch.n_accepted[b]) == a.n_accepted and int(batch.n_rejected[b]) == a.n_rejected
        assert int(batch.status[b]) == a.status
        assert torch.equal(batch.dt_history[b], a.dt_history)
        for f in ("u", "v"):
            assert torch.equal(batch.y_final[f][b], a.y_final[f])
            assert torch.equal(batch.y_history[f][b], a.y_history[f])


def test_adaptive_fp32_state(gpu):
    ref, res, _ = _run_both(64, 1, 20.0, 0.5, dtype=torch.float32)
    seq_ref = [a[3] for a in ref.attempts]
    print(f"fp32 state: acc={res.n_accepted} rej={res.n_rejected} vs fp64 oracle acc={ref.n_accepted} "
          f"rej={ref.n_rejected}", flush=True)
    assert res.status == oad.SUCCESS and res.t_final == pytest.approx(20.0, abs=1e-12)
    assert abs(res.n_accepted - ref.n_accepted) <= 2 and abs(res.n_rejected - ref.n_rejected) <= 2
    assert len(seq_ref) == ref.n_accepted + ref.n_rejected
    assert ogs.rel_l2(_np(res.y_final["v"]), ref.y_final["v"]) < 1e-3


def test_adaptive_status_codes(gpu):
    from moljax_amd.core.dt_policy import PIDParams
    from moljax_amd.core.stepping import adaptive_integrate_imex
    n = 32
    (model, cache, diff), p = _model(n, torch.float64)
    y0 = ogs.initial_condition(n, n, seed=0, square=6)
    r = adaptive_integrate_imex(model, _dev(y0, torch.float64), 1.0, 1.0, 0.1, cache, diff)
    assert r.status == oad.SUCCESS and r.n_accepted == 0 and r.n_steps == 1
    r = adaptive_integrate_imex(model, _dev(y0, torch.float64), 0.0, 100.0, 0.1, cache, diff, max_steps=3)
    assert r.status == oad.MAX_STEPS_REACHED and r.n_steps == 4
    ref = oad.adaptive_integrate_imex(y0, 0.0, 1.0, 0.5, p["F"], p["k"], p["Du"], p["Dv"], _lam(n), L / n,
                                      pid=oad.PIDParams(atol=1e-12, rtol=1e-12, dt_min=0.05))
    r = adaptive_integrate_imex(model, _dev(y0, torch.float64), 0.0, 1.0, 0.5, cache, diff,
                                pid_params=PIDParams(atol=1e-12, rtol=1e-12, dt_min=0.05))
    assert r.status == ref.status == oad.DT_TOO_SMALL and r.n_rejected == ref.n_rejected
    bad = {k: v.copy() for k, v in y0.items()}
    bad["u"][5, 5] = np.nan
    r = adaptive_integrate_imex(model, _dev(bad, torch.float64), 0.0, 1.0, 0.1, cache, diff)
    assert r.status == oad.NON_FINITE_VALUES and r.n_accepted == 0 and r.n_rejected == 1


def test_adaptive_max_attempts_resumes_as_oracle(gpu):
    from moljax_amd.core.dt_policy import PIDParams
    from moljax_amd.core.stepping import adaptive_integrate_imex
    n = 32
    (model, cache, diff), p = _model(n, torch.float64)
    y0 = ogs.initial_condition(n, n, seed=5, square=8)
    ref = oad.adaptive_integrate_imex(y0, 0.0, 10.0, 0.5, p["F"], p["k"], p["Du"], p["Dv"], _lam(n), L / n,
                                      pid=oad.PIDParams(**PID), max_attempts=7)
    r = adaptive_integrate_imex(model, _dev(y0, torch.float64), 0.0, 10.0, 0.5, cache, diff,
                                pid_params=PIDParams(**PID), max_attempts=7)
    assert r.status == oad.RUNNING == ref.status
    assert (r.n_accepted, r.n_rejected) == (ref.n_accepted, ref.n_rejected)
    assert r.t_final == pytest.approx(ref.t_final, rel=1e-12)
