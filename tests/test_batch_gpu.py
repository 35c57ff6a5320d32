"""Batched uneven-size plans (hipGraph replay, autotuned tiles) on the GPU.

Each instance of a mixed-size batch (power-of-two and other sides) is
bitwise equal to the same instance run alone (a plan of one) and matches the
float64 oracle (ETDRK4 <= 1e-12,
IMEX-Strang <= 1e-12 relative L2 after several steps)."""
import numpy as np
import pytest
import torch

from oracle import etdrk4 as oetd
from oracle import gray_scott as ogs

pytestmark = pytest.mark.gpu

DU, DV, DT = 2e-5, 1e-5, 1.0


def _ic(n, seed):
    from moljax_amd.workloads import gray_scott_ic
    u, v = gray_scott_ic(n, n, seed, square=max(4, n // 8))
    return u[1:-1, 1:-1].copy(), v[1:-1, 1:-1].copy()


def _batch(sizes, F, k, scheme, dtype, steps_per_graph=2):
    from moljax_amd.batch import GrayScottBatch
    b = GrayScottBatch(sizes, F, k, DU, DV, DT, scheme=scheme, dtype=dtype, steps_per_graph=steps_per_graph)
    for i, n in enumerate(sizes):
        u, v = _ic(n, seed=i)
        b.set_state(i, torch.from_numpy(u), torch.from_numpy(v))
    return b


@pytest.mark.parametrize("scheme", ["etdrk4", "strang"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("sizes", [[64, 32, 64, 128, 32, 64], [48, 64, 100, 48, 33]])
def test_mixed_batch_equals_alone_and_oracle(gpu, scheme, dtype, sizes):
    """Power-of-two groups (Stockham kernels) and, in the second size set,
    groups of other sides (chirp-z kernels) in one plan."""
    from moljax_amd.batch import GrayScottBatch, spectral_symbol_1d
    F = [0.02 + 0.005 * i for i in range(len(sizes))]
    k = [0.055 + 0.002 * i for i in range(len(sizes))]
    steps = 5
    mix = _batch(sizes, F, k, scheme, dtype)
    mix.run(steps)
    for i, n in enumerate(sizes):
        alone = GrayScottBatch([n], [F[i]], [k[i]], DU, DV, DT, scheme=scheme, dtype=dtype, steps_per_graph=2)
        u0, v0 = _ic(n, seed=i)
        alone.set_state(0, torch.from_numpy(u0), torch.from_numpy(v0))
        alone.run(steps)
        mu, mv = mix.state(i)
        au, av = alone.state(0)
        assert torch.equal(mu, au) and torch.equal(mv, av), f"instance {i} (n={n}) differs from its lone run"
        alone.close()
        if dtype != torch.float64:
            continue
        lam = spectral_symbol_1d(n)
        if scheme == "etdrk4":
            ru, rv = oetd.gs_etdrk4(u0, v0, lam, lam, DU, DV, F[i], k[i], DT, steps)
        else:
            P = lambda a: ogs.apply_bc_periodic(np.pad(a, 1))
            y = {"u": P(u0), "v": P(v0)}
            lam2 = lam[:, None] + lam[None, :]
            for _ in range(steps):
                y = ogs.imex_strang_step(y, DT, F[i], k[i], DU, DV, lam2)
            ru, rv = y["u"][1:-1, 1:-1], y["v"][1:-1, 1:-1]
        assert ogs.rel_l2(mu.cpu().numpy(), ru) < 1e-12
        assert ogs.rel_l2(mv.cpu().numpy(), rv) < 1e-12
    mix.close()


def test_graph_replay_count_independent(gpu):
    """The same run split over different graph lengths gives the same bits."""
    sizes = [32, 64, 32]
    F, k = [0.03, 0.04, 0.05], [0.06, 0.062, 0.064]
    outs = []
    for spg in (1, 3, 16):
        b = _batch(sizes, F, k, "etdrk4", torch.float32, steps_per_graph=spg)
        b.run(7)
        outs.append([t.clone() for i in range(3) for t in b.state(i)])
        b.close()
    for o in outs[1:]:
        assert all(torch.equal(a, c) for a, c in zip(outs[0], o))


def test_autotune_and_tile_table(gpu):
    from moljax_amd import _hip
    from moljax_amd.batch import autotune
    lib = _hip.lib()
    for scheme in (0, 1):
        r, c = autotune(torch.float32, scheme, 64, 8)
        print(f"scheme {scheme} n=64: row lines {r}, col lines {c}", flush=True)
        assert r >= 1 and c >= 1 and (r & (r - 1)) == 0 and (c & (c - 1)) == 0
        kind_r = 2 if scheme == 0 else 0
        assert lib.mjx_tune_get(kind_r, 0, 64) == r
    assert lib.mjx_tune_set(0, 0, 128, 4) == 0 and lib.mjx_tune_get(0, 0, 128) == 4
    assert lib.mjx_tune_set(0, 0, 128, 3) != 0  # not a power of two
    assert lib.mjx_tune_set(0, 0, 128, 0) == 0  # back to the default


@pytest.mark.parametrize("n", [128, 512, 1024])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 1e-5)])
def test_etdrk4_every_column_tile_matches_oracle(gpu, n, dtype, tol):
    """Each ETDRK4 column tile the autotuner may choose (widest that fits a
    workgroup, /2, /4, /8) gives the oracle's result (3 steps).  FP64 runs
    stage 3 as two kernels, FP32 as the merged one."""
    from moljax_amd import _hip
    from moljax_amd.batch import spectral_symbol_1d
    from moljax_amd.core.fft_integrators import etdrk4_coefficients, gs_etdrk4_inplace
    lib = _hip.lib()
    code = _hip.dtype_code(dtype)
    lam = spectral_symbol_1d(n)
    u0, v0 = _ic(n, 3)
    F, k, steps = 0.04, 0.06, 3
    ru, rv = oetd.gs_etdrk4(u0, v0, lam, lam, DU, DV, F, k, DT, steps)
    coef = etdrk4_coefficients(lam, lam, DU, DV, DT, dtype)
    tn = n // {128: 16, 512: 8, 1024: 16}[n]
    es = 16 if dtype == torch.float64 else 8  # bytes of one complex value
    cm = min(1024 // tn, n // 2, 65536 // (n * es))
    try:
        for L in sorted({cm, cm // 2, cm // 4, cm // 8} - {0}):
            assert lib.mjx_tune_set(3, code, n, L) == 0
            U, V = torch.from_numpy(u0).to("cuda", dtype), torch.from_numpy(v0).to("cuda", dtype)
            gs_etdrk4_inplace(U, V, coef, [[0, 0, 0, 0, F, k, DT]], steps)
            eu, ev = ogs.rel_l2(U.double().cpu().numpy(), ru), ogs.rel_l2(V.double().cpu().numpy(), rv)
            print(f"n={n} {dtype} column tile {L}: rel_l2 u={eu:.1e} v={ev:.1e}", flush=True)
            assert eu < tol and ev < tol
    finally:
        lib.mjx_tune_set(3, code, n, 0)


def test_strang_every_tile_matches_oracle(gpu):
    """Each row / column tile of the Strang chain (lines per workgroup from 1
    up to the default) gives the oracle's result (FP64, 256^2, 3 steps)."""
    from moljax_amd import _hip
    from moljax_amd.batch import fd_symbol_1d
    from moljax_amd.core._arrays import inst_table, spectral_work
    lib = _hip.lib()
    n, steps, F, k = 256, 3, 0.035, 0.065
    lam = fd_symbol_1d(n, 2.5)
    u0, v0 = _ic(n, 4)
    P = lambda a: ogs.apply_bc_periodic(np.pad(a, 1))
    y = {"u": P(u0), "v": P(v0)}
    for _ in range(steps):
        y = ogs.imex_strang_step(y, DT, F, k, DU, DV, lam[:, None] + lam[None, :])
    ref_u = y["u"][1:-1, 1:-1]
    lt = torch.from_numpy(lam).cuda()
    rows = inst_table([[0.5 * DT * DU, 0.5 * DT * DV, 0, 0, F, k, DT]])
    work = spectral_work(torch.float64, 1, n, n)
    try:
        for kind in (0, 1):
            for L in (1, 2, 4, 8):
                lib.mjx_tune_set(0, 1, n, 0)
                lib.mjx_tune_set(1, 1, n, 0)
                assert lib.mjx_tune_set(kind, 1, n, L) == 0
                U, V = torch.from_numpy(u0).cuda(), torch.from_numpy(v0).cuda()
                rc = lib.mjx_gs_imex_strang(1, 1, n, n, _hip.ptr(U), _hip.ptr(V), n, n * n, 0, _hip.ptr(lt),
                                            _hip.ptr(lt), None, _hip.ptr(rows), steps, _hip.ptr(work),
                                            _hip.stream_ptr())
                _hip.check(rc, "strang")
                err = ogs.rel_l2(U.cpu().numpy(), ref_u)
                print(f"{'row' if kind == 0 else 'col'} tile {L}: rel_l2={err:.1e}", flush=True)
                assert err < 1e-12
    finally:
        lib.mjx_tune_set(0, 1, n, 0)
        lib.mjx_tune_set(1, 1, n, 0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_c4_four_size_plan(gpu, dtype):
    """The batched ETDRK4 workload's own shape: one plan holding all four
    sizes {64, 128, 256, 512}^2 (two instances each, c4's F, k ranges and IC
    convention), 3 steps.  Each instance is bitwise equal to its lone run;
    in FP64 each is within 1e-12 of the oracle, in FP32 within 1e-5."""
    from moljax_amd.batch import GrayScottBatch, spectral_symbol_1d
    from moljax_amd.workloads import gray_scott_ic
    sizes = [512, 64, 256, 128, 64, 512, 128, 256]
    rng = np.random.default_rng(11)
    F = rng.uniform(0.02, 0.06, len(sizes)).tolist()
    k = rng.uniform(0.055, 0.068, len(sizes)).tolist()
    ics = []
    for i, n in enumerate(sizes):
        u, v = gray_scott_ic(n, n, i, square=max(2, round(20 * n / 256)))
        ics.append((u[1:-1, 1:-1].copy(), v[1:-1, 1:-1].copy()))
    steps = 3
    mix = GrayScottBatch(sizes, F, k, DU, DV, DT, scheme="etdrk4", dtype=dtype, steps_per_graph=2)
    for i, (u, v) in enumerate(ics):
        mix.set_state(i, torch.from_numpy(u), torch.from_numpy(v))
    mix.run(steps)
    for i, n in enumerate(sizes):
        print(f"c4 plan instance {i} n={n} {dtype}", flush=True)
        alone = GrayScottBatch([n], [F[i]], [k[i]], DU, DV, DT, scheme="etdrk4", dtype=dtype, steps_per_graph=2)
        alone.set_state(0, torch.from_numpy(ics[i][0]), torch.from_numpy(ics[i][1]))
        alone.run(steps)
        mu, mv = mix.state(i)
        au, av = alone.state(0)
        assert torch.equal(mu, au) and torch.equal(mv, av), f"instance {i} (n={n}) differs from its lone run"
        alone.close()
        lam = spectral_symbol_1d(n)
        ru, rv = oetd.gs_etdrk4(ics[i][0], ics[i][1], lam, lam, DU, DV, F[i], k[i], DT, steps)
        tol = 1e-12 if dtype == torch.float64 else 1e-5
        assert ogs.rel_l2(mu.double().cpu().numpy(), ru) < tol
        assert ogs.rel_l2(mv.double().cpu().numpy(), rv) < tol
    mix.close()


def _helm_oracle(kind, u, n, D, dt, steps):
    from oracle import nonperiodic as onp
    from oracle import spectral as osp
    for _ in range(steps):
        if kind == "periodic":
            u = osp.solve_helmholtz_2d(u, osp.laplacian_symbol_2d(n, n, 1.0 / n, 1.0 / n), dt, D)
        elif kind == "dirichlet":
            lam = onp.laplacian_symbol_dirichlet(n, 1.0 / (n + 1))
            u = onp.solve_helmholtz_2d("dirichlet", u, lam, lam, dt, D)
        else:
            lam = onp.laplacian_symbol_neumann(n, 1.0 / n)
            u = onp.solve_helmholtz_2d("neumann", u, lam, lam, dt, D)
    return u


def _mixed(dtype, spg, helm_only=None):
    """Gray-Scott (ETDRK4, Strang) and Helmholtz (periodic, Dirichlet,
    Neumann) instances of uneven power-of-two and other sizes."""
    from moljax_amd.batch import MixedBatch
    mb = MixedBatch(dtype=dtype, steps_per_graph=spg)
    specs = []
    gs = [("etdrk4", [64, 48]), ("strang", [32])]
    helm = [("periodic", [64, 40]), ("dirichlet", [63, 50]), ("neumann", [64, 37])]
    rng = np.random.default_rng(5)
    for scheme, sizes in gs:
        F = [0.03 + 0.004 * i for i in range(len(sizes))]
        k = [0.06 + 0.001 * i for i in range(len(sizes))]
        ids = mb.add_gray_scott(sizes, F, k, DU, DV, DT, scheme=scheme)
        for i, n, f, kk in zip(ids, sizes, F, k):
            specs.append(("gs", scheme, n, dict(F=f, k=kk), _ic(n, seed=i)))
    for kind, sizes in helm:
        D = [10 ** rng.uniform(-3, -1) for _ in sizes]
        dt = [10 ** rng.uniform(-3, -1) for _ in sizes]
        ids = mb.add_helmholtz(kind, sizes, D, dt)
        for i, n, d, t in zip(ids, sizes, D, dt):
            specs.append(("helm", kind, n, dict(D=d, dt=t), (rng.standard_normal((n, n)),)))
    mb.build()
    for i, s in enumerate(specs):
        mb.set_state(i, *[torch.from_numpy(a) for a in s[4]])
    return mb, specs


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_mixed_kinds_one_plan(gpu, dtype):
    """Gray-Scott and Helmholtz groups of uneven sizes in one plan and one
    graph: each instance bitwise equal to the same instance in a plan of its
    own; in FP64 each Helmholtz instance within 1e-12 of the oracle's
    step-by-step solves and each Gray-Scott instance within 1e-12 of the
    oracle."""
    from moljax_amd.batch import MixedBatch, spectral_symbol_1d
    steps = 5
    mb, specs = _mixed(dtype, spg=2)
    mb.run(steps)
    for i, (fam, kind, n, prm, ic) in enumerate(specs):
        print(f"mixed plan instance {i}: {fam} {kind} n={n} {dtype}", flush=True)
        alone = MixedBatch(dtype=dtype, steps_per_graph=2)
        if fam == "gs":
            alone.add_gray_scott([n], [prm["F"]], [prm["k"]], DU, DV, DT, scheme=kind)
        else:
            alone.add_helmholtz(kind, [n], prm["D"], prm["dt"])
        alone.build()
        alone.set_state(0, *[torch.from_numpy(a) for a in ic])
        alone.run(steps)
        for a, b in zip(mb.state(i), alone.state(0)):
            assert torch.equal(a, b), f"instance {i} ({fam} {kind}, n={n}) differs from its lone run"
        alone.close()
        if dtype != torch.float64:
            continue
        got = [t.cpu().numpy() for t in mb.state(i)]
        if fam == "helm":
            ref = [_helm_oracle(kind, ic[0], n, prm["D"], prm["dt"], steps)]
        elif kind == "etdrk4":
            lam = spectral_symbol_1d(n)
            ref = list(oetd.gs_etdrk4(ic[0], ic[1], lam, lam, DU, DV, prm["F"], prm["k"], DT, steps))
        else:
            lam = spectral_symbol_1d(n)
            P = lambda a: ogs.apply_bc_periodic(np.pad(a, 1))
            y = {"u": P(ic[0]), "v": P(ic[1])}
            for _ in range(steps):
                y = ogs.imex_strang_step(y, DT, prm["F"], prm["k"], DU, DV, lam[:, None] + lam[None, :])
            ref = [y["u"][1:-1, 1:-1], y["v"][1:-1, 1:-1]]
        for g, r in zip(got, ref):
            err = ogs.rel_l2(g, r)
            assert err < 1e-12, (fam, kind, n, err)
    mb.close()


def test_mixed_plan_graph_length_independent(gpu):
    """The same mixed plan run with different graph lengths gives the same
    bits (the Strang groups' middle steps differ from the others' by one)."""
    outs = []
    for spg in (1, 3, 8):
        mb, specs = _mixed(torch.float32, spg)
        mb.run(7)
        outs.append([t.clone() for i in range(len(specs)) for t in mb.state(i)])
        mb.close()
    for o in outs[1:]:
        assert all(torch.equal(a, c) for a, c in zip(outs[0], o))
