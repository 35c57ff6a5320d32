"""Continuous-batching request loop on the GPU.

* alone == in the mix, bitwise, for requests of every kind;
* one request per kind (float64) against the float64 oracle (<= 1e-11
  relative L2): FFT-CN, DST-I and DCT-II backward-Euler diffusion, and
  Gray-Scott IMEX-Strang;
* admission by HBM footprint holds the active set within the budget."""
import numpy as np
import pytest
import torch

from oracle import gray_scott as ogs
from oracle import nonperiodic as onp
from oracle import spectral as osp

pytestmark = pytest.mark.gpu


def _mix():
    from moljax_amd.serving import Request
    reqs, rid = [], 0
    for kind in ("fft_cn", "dst", "dct", "imex_gs"):
        for n, steps in ((32, 37), (64, 70), (32, 40), (32, 5)):
            reqs.append(Request(rid=rid, kind=kind, n=n, D=1e-2 * (1 + rid % 3), dt=2e-2 * (1 + rid % 2),
                                steps=steps, seed=100 + rid, arrival=0.0))
            rid += 1
    return reqs


def _fresh(r):
    from moljax_amd.serving import Request
    return Request(rid=r.rid, kind=r.kind, n=r.n, D=r.D, dt=r.dt, steps=r.steps, seed=r.seed, arrival=0.0)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_alone_equals_mix(gpu, dtype):
    from moljax_amd.serving import Server
    mix = _mix()
    srv = Server(dtype=dtype, chunk=16)
    srv.run(mix, realtime=False)
    print(f"{dtype}: iterations={srv.iterations} peak_active={srv.peak_active} max_group={srv.max_group}", flush=True)
    assert srv.max_group >= 2  # requests really shared launches
    for r in mix:
        alone = _fresh(r)
        Server(dtype=dtype, chunk=16).run([alone], realtime=False)
        for a, b in zip(r.state, alone.state):
            assert torch.equal(a, b), f"request {r.rid} ({r.kind}, n={r.n}) differs alone vs in the mix"


def _oracle(r, fields):
    m = r.grid
    if r.kind == "fft_cn":
        G = osp.cn_factor(osp.spectral_symbol_ij(m).T, r.D, r.dt)
        u = fields[0]
        for _ in range(r.steps):
            u = np.real(np.fft.ifft2(np.fft.fft2(u) * G))
        return [u]
    if r.kind in ("dst", "dct"):
        if r.kind == "dst":
            lam = onp.laplacian_symbol_dirichlet(m, 1.0 / (m + 1))
            kind = "dirichlet"
        else:
            lam = onp.laplacian_symbol_neumann(m, 1.0 / m)
            kind = "neumann"
        u = fields[0]
        for _ in range(r.steps):
            u = onp.solve_helmholtz_2d(kind, u, lam, lam, r.dt, r.D)
        return [u]
    from moljax_amd.serving import GS_F, GS_K, GS_L
    lam = osp.laplacian_symbol_2d(m, m, GS_L / m, GS_L / m)
    P = lambda a: ogs.apply_bc_periodic(np.pad(a, 1))
    y = {"u": P(fields[0]), "v": P(fields[1])}
    for _ in range(r.steps):
        y = ogs.imex_strang_step(y, r.dt, GS_F, GS_K, r.D, r.D / 2, lam)
    return [y["u"][1:-1, 1:-1], y["v"][1:-1, 1:-1]]


def test_each_kind_matches_oracle(gpu):
    from moljax_amd.serving import Server, initial_state
    reqs = [r for r in _mix() if r.n == 64]
    Server(dtype=torch.float64, chunk=16).run(reqs, realtime=False)
    for r in reqs:
        ic = [t.cpu().numpy() for t in initial_state(_fresh(r), torch.float64)]
        ref = _oracle(r, ic)
        for got, want in zip(r.state, ref):
            err = ogs.rel_l2(got.cpu().numpy(), want)
            print(f"{r.kind} n={r.n} steps={r.steps}: rel_l2={err:.2e}", flush=True)
            assert err < 1e-11, r.kind


def test_admission_by_footprint(gpu):
    from moljax_amd.serving import Server, latency_stats, poisson_stream
    reqs = poisson_stream(16, rate=2000.0, seed=3, sizes=(32, 64), steps=(5, 40))
    probe = Server()
    budget = 2 * max(probe.footprint(r) for r in reqs)
    srv = Server(hbm_budget=budget, chunk=8)
    srv.run(reqs, realtime=True)
    assert srv.peak_used <= budget and srv.peak_active >= 1
    assert all(np.isfinite(r.finished) and r.finished >= r.admitted >= 0 for r in reqs)
    st = latency_stats(reqs)
    assert st["p99_ms"] >= st["p50_ms"] > 0
    again = poisson_stream(16, rate=2000.0, seed=3, sizes=(32, 64), steps=(5, 40))
    assert [(a.kind, a.n, a.steps, a.seed) for a in again] == [(a.kind, a.n, a.steps, a.seed) for a in reqs]


def test_masked_chain_equals_fixed_chains(gpu):
    """mjx_gs_imex_strang_var: each instance's result is bitwise that of the
    fixed-length chain with its own step count."""
    from moljax_amd import _hip
    from moljax_amd.core._arrays import inst_table, spectral_work
    from moljax_amd.serving import _fd, GS_L
    m, steps = 64, [3, 5, 1, 4]
    B = len(steps)
    lib = _hip.lib()
    lam = torch.from_numpy(_fd(m, GS_L)).cuda()
    rng = np.random.default_rng(0)
    u0 = torch.from_numpy(1.0 - 0.5 * rng.random((B, m, m))).cuda()
    v0 = torch.from_numpy(0.25 * rng.random((B, m, m))).cuda()
    rows = [[0.5 * 0.3 * 1e-3 * (b + 1), 0.25 * 0.3 * 1e-3, 0, 0, 0.035, 0.065, 0.3, steps[b]] for b in range(B)]
    u, v = u0.clone(), v0.clone()
    # every device buffer stays referenced until its kernels have run
    inst, work = inst_table(rows), spectral_work(torch.float64, B, m, m)
    rc = lib.mjx_gs_imex_strang_var(1, B, m, m, _hip.ptr(u), _hip.ptr(v), m, m * m, 0, _hip.ptr(lam), _hip.ptr(lam),
                                    None, _hip.ptr(inst), max(steps), _hip.ptr(work), _hip.stream_ptr())
    _hip.check(rc, "var")
    for b in range(B):
        ub, vb = u0[b:b + 1].clone(), v0[b:b + 1].clone()
        ib, wb = inst_table([rows[b]]), spectral_work(torch.float64, 1, m, m)
        rc = lib.mjx_gs_imex_strang(1, 1, m, m, _hip.ptr(ub), _hip.ptr(vb), m, m * m, 0, _hip.ptr(lam), _hip.ptr(lam),
                                    None, _hip.ptr(ib), steps[b], _hip.ptr(wb), _hip.stream_ptr())
        _hip.check(rc, "fixed")
        du = float((u[b] - ub[0]).abs().max())
        dv = float((v[b] - vb[0]).abs().max())
        print(f"instance {b} steps {steps[b]}: max |du| {du:.3e} |dv| {dv:.3e}", flush=True)
        assert torch.equal(u[b], ub[0]) and torch.equal(v[b], vb[0]), b


def test_any_length_requests(gpu):
    """Sides outside the power-of-two rule (the chirp-z kernels) served in
    one mix with power-of-two sides: alone == mix bitwise (float32), and each
    kind against the oracle (float64, <= 1e-11)."""
    from moljax_amd.serving import Request, Server, initial_state
    specs = []
    for kind in ("fft_cn", "dst", "dct", "imex_gs"):
        for n, steps in ((48, 13), (40, 9), (64, 11)):
            specs.append((kind, n, steps))
    mk = lambda: [Request(rid=i, kind=k, n=n, D=1e-2 * (1 + i % 3), dt=2e-2, steps=s, seed=300 + i, arrival=0.0)
                  for i, (k, n, s) in enumerate(specs)]
    mix = mk()
    Server(dtype=torch.float32, chunk=8).run(mix, realtime=False)
    for r in mix:
        alone = _fresh(r)
        Server(dtype=torch.float32, chunk=8).run([alone], realtime=False)
        for a, b in zip(r.state, alone.state):
            assert torch.equal(a, b), f"request {r.rid} ({r.kind}, n={r.n}) differs alone vs in the mix"
    reqs = [r for r in mk() if r.n != 64]
    Server(dtype=torch.float64, chunk=8).run(reqs, realtime=False)
    for r in reqs:
        ic = [t.cpu().numpy() for t in initial_state(_fresh(r), torch.float64)]
        for got, want in zip(r.state, _oracle(r, ic)):
            assert ogs.rel_l2(got.cpu().numpy(), want) < 1e-11, (r.kind, r.n)


def test_c5_sizes_up_to_1024(gpu):
    """The serving workload's own size range: requests of every kind at
    1024^2 beside 64^2 and 512^2 ones, served together.  Alone == mix
    bitwise (float32); each 1024^2 kind against the oracle (float64,
    <= 1e-11 relative L2)."""
    from moljax_amd.serving import Request, Server, initial_state
    specs = []
    for kind in ("fft_cn", "dst", "dct", "imex_gs"):
        for n, steps in ((1024, 9), (64, 13), (512, 6)):
            specs.append((kind, n, steps))
    mk = lambda: [Request(rid=i, kind=k, n=n, D=10 ** (-1 - (i % 3)), dt=10 ** (-1 - (i % 2)), steps=s,
                          seed=700 + i, arrival=0.0) for i, (k, n, s) in enumerate(specs)]
    mix = mk()
    Server(dtype=torch.float32, chunk=4).run(mix, realtime=False)
    for r in mix:
        print(f"c5 alone vs mix: {r.kind} n={r.n}", flush=True)
        alone = _fresh(r)
        Server(dtype=torch.float32, chunk=4).run([alone], realtime=False)
        for a, b in zip(r.state, alone.state):
            assert torch.equal(a, b), f"request {r.rid} ({r.kind}, n={r.n}) differs alone vs in the mix"
    reqs = [r for r in mk() if r.n == 1024]
    Server(dtype=torch.float64, chunk=4).run(reqs, realtime=False)
    for r in reqs:
        ic = [t.cpu().numpy() for t in initial_state(_fresh(r), torch.float64)]
        for got, want in zip(r.state, _oracle(r, ic)):
            err = ogs.rel_l2(got.cpu().numpy(), want)
            print(f"c5 {r.kind} n={r.n} steps={r.steps}: rel_l2={err:.2e}", flush=True)
            assert err < 1e-11, (r.kind, r.n)
