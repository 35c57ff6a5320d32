"""Canary-filled outputs: every output cell an entry point must write is
pre-filled with NaN and every cell it must not touch (the pitch padding
beyond the padded row, a spare instance after the batch) holds a different
canary.  After the call, the written region is finite and the untouched
region still holds its canary bit for bit, so an unwritten ghost or
interior cell, or a write past the described layout, fails the test."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

CANARY = 12345.678


def _buffer(B, rows, pitch, dtype, fill=float("nan")):
    """B + 1 instances of rows x pitch, NaN everywhere; the spare instance
    and the columns past the used width get the canary value later."""
    return torch.full((B + 1, rows, pitch), fill, dtype=dtype, device="cuda")


def _mark_outside(buf, B, width):
    buf[:, :, width:] = CANARY
    buf[B:] = CANARY


def _check(buf, B, rows_used, width, name):
    used = buf[:B, :rows_used, :width]
    assert torch.isfinite(used).all(), f"{name}: an output cell was not written"
    assert (buf[:, :, width:] == CANARY).all(), f"{name}: wrote past the padded row"
    assert (buf[B:] == CANARY).all(), f"{name}: wrote past the last instance"


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("op", [0, 1, 2, 3])
def test_fd_operator_writes_interior_and_ghost_ring(gpu, dtype, op):
    from moljax_amd import _hip
    B, ny, nx, P = 3, 37, 50, 56
    W = nx + 2
    src = torch.randn((B, ny + 2, P), dtype=dtype, device="cuda")
    out = _buffer(B, ny + 2, P, dtype)
    _mark_outside(out, B, W)
    rc = _hip.lib().mjx_fd_op2d(_hip.dtype_code(dtype), op, B, ny, nx, _hip.ptr(src[:, 1:, 1:]), P, (ny + 2) * P,
                                _hip.ptr(out[:, 1:, 1:]), P, (ny + 2) * P, op % 2, 0.1, 0.2, None, 0, 0, -0.5,
                                _hip.stream_ptr())
    _hip.check(rc, "fd_op2d")
    torch.cuda.synchronize()
    _check(out, B, ny + 2, W, f"fd op {op}")
    ring = torch.cat([out[:B, 0, :W], out[:B, -1, :W], out[:B, :, 0], out[:B, :, W - 1]], dim=-1)
    assert (ring == 0).all(), "output ghost ring not zero"


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_bc_fill_writes_the_whole_ring(gpu, kind):
    from moljax_amd import _hip
    import ctypes
    B, ny, nx, P, dtype = 2, 33, 45, 52, torch.float64
    W = nx + 2
    f = _buffer(B, ny + 2, P, dtype)
    f[:B, 1:-1, 1:W - 1] = torch.randn((B, ny, nx), dtype=dtype, device="cuda")
    _mark_outside(f, B, W)
    vals = (ctypes.c_double * 4)(0.1, -0.2, 0.3, 0.4)
    rc = _hip.lib().mjx_bc_fill2d(1, B, ny, nx, _hip.ptr(f[:, 1:, 1:]), P, (ny + 2) * P, kind, vals, 0.05, 0.04,
                                  _hip.stream_ptr())
    _hip.check(rc, "bc_fill2d")
    torch.cuda.synchronize()
    _check(f, B, ny + 2, W, f"bc fill kind {kind}")


@pytest.mark.parametrize("n", [64, 48])
def test_imex_strang_with_ghosts_writes_every_padded_cell(gpu, n):
    from moljax_amd import _hip
    from moljax_amd.batch import fd_symbol_1d
    from moljax_amd.core._arrays import inst_table, spectral_work
    B, P, dtype = 2, n + 8, torch.float32
    W = n + 2
    u = _buffer(B, n + 2, P, dtype)
    v = _buffer(B, n + 2, P, dtype)
    rng = np.random.default_rng(0)
    u[:B, 1:-1, 1:W - 1] = torch.from_numpy(1.0 - 0.3 * rng.random((B, n, n))).to("cuda", dtype)
    v[:B, 1:-1, 1:W - 1] = torch.from_numpy(0.2 * rng.random((B, n, n))).to("cuda", dtype)
    for t in (u, v):
        _mark_outside(t, B, W)
    lam = torch.from_numpy(fd_symbol_1d(n, 2.5)).to("cuda", dtype)
    rows = inst_table([[1e-5, 5e-6, 0, 0, 0.035, 0.065, 1.0]] * B)
    work = spectral_work(dtype, B, n, n)
    rc = _hip.lib().mjx_gs_imex_strang(0, B, n, n, _hip.ptr(u[:, 1:, 1:]), _hip.ptr(v[:, 1:, 1:]), P, (n + 2) * P, 1,
                                       _hip.ptr(lam), _hip.ptr(lam), None, _hip.ptr(rows), 3, _hip.ptr(work),
                                       _hip.stream_ptr())
    _hip.check(rc, "gs_imex_strang")
    torch.cuda.synchronize()
    _check(u, B, n + 2, W, "strang u")
    _check(v, B, n + 2, W, "strang v")


@pytest.mark.parametrize("kind,n", [(0, 63), (1, 64), (0, 40), (1, 37)])
def test_r2r_solve_writes_the_interior_only(gpu, kind, n):
    from moljax_amd import _hip
    from moljax_amd.batch import dirichlet_symbol_1d, neumann_symbol_1d
    from moljax_amd.core._arrays import inst_table
    B, P, dtype = 3, n + 5, torch.float64
    rhs = torch.randn((B, n, P), dtype=dtype, device="cuda")
    out = _buffer(B, n, P, dtype)
    _mark_outside(out, B, n)
    lam = torch.from_numpy((dirichlet_symbol_1d if kind == 0 else neumann_symbol_1d)(n)).to("cuda", dtype)
    rows = inst_table([[0.01, 0, 0, 0, 0, 0, 0.1]] * B)
    work = torch.empty(B * n * n, dtype=dtype, device="cuda")
    rc = _hip.lib().mjx_r2r_solve2d(1, kind, B, n, n, _hip.ptr(rhs), P, n * P, _hip.ptr(out), P, n * P,
                                    _hip.ptr(lam), _hip.ptr(lam), _hip.ptr(rows), 0, _hip.ptr(work),
                                    _hip.stream_ptr())
    _hip.check(rc, "r2r_solve2d")
    torch.cuda.synchronize()
    _check(out, B, n, n, f"r2r kind {kind} n={n}")


@pytest.mark.parametrize("n", [64, 50])
def test_etdrk4_with_ghosts_writes_every_padded_cell(gpu, n):
    from moljax_amd import _hip
    from moljax_amd.batch import spectral_symbol_1d
    from moljax_amd.core._arrays import inst_table
    from moljax_amd.core.fft_integrators import etdrk4_coefficients
    B, P, dtype = 2, n + 6, torch.float32
    W = n + 2
    u = _buffer(B, n + 2, P, dtype)
    v = _buffer(B, n + 2, P, dtype)
    rng = np.random.default_rng(1)
    u[:B, 1:-1, 1:W - 1] = torch.from_numpy(1.0 - 0.3 * rng.random((B, n, n))).to("cuda", dtype)
    v[:B, 1:-1, 1:W - 1] = torch.from_numpy(0.2 * rng.random((B, n, n))).to("cuda", dtype)
    for t in (u, v):
        _mark_outside(t, B, W)
    lam = spectral_symbol_1d(n)
    coef = etdrk4_coefficients(lam, lam, 2e-5, 1e-5, 1.0, dtype)
    rows = inst_table([[0, 0, 0, 0, 0.04, 0.06, 1.0]] * B)
    work = torch.empty(_hip.lib().mjx_etdrk4_workspace_bytes(0, B, n, n), dtype=torch.uint8, device="cuda")
    rc = _hip.lib().mjx_gs_etdrk4(0, B, n, n, _hip.ptr(u[:, 1:, 1:]), _hip.ptr(v[:, 1:, 1:]), P, (n + 2) * P, 1,
                                  _hip.ptr(coef), _hip.ptr(rows), 2, _hip.ptr(work), _hip.stream_ptr())
    _hip.check(rc, "gs_etdrk4")
    torch.cuda.synchronize()
    _check(u, B, n + 2, W, "etdrk4 u")
    _check(v, B, n + 2, W, "etdrk4 v")


def test_reduction_writes_every_instance(gpu):
    from moljax_amd import _hip
    B, ny, nx, P, dtype = 5, 70, 33, 40, torch.float32
    a = torch.randn((B, ny, P), dtype=dtype, device="cuda")
    out = torch.full((B + 1,), float("nan"), dtype=torch.float64, device="cuda")
    out[B] = CANARY
    work = torch.empty(_hip.lib().mjx_reduce_workspace_elems(B), dtype=torch.float64, device="cuda")
    rc = _hip.lib().mjx_reduce2d(0, 0, B, ny, nx, _hip.ptr(a), P, ny * P, None, 0, 0, 0.0, 0.0, _hip.ptr(out), 0,
                                 _hip.ptr(work), _hip.stream_ptr())
    _hip.check(rc, "reduce2d")
    torch.cuda.synchronize()
    assert torch.isfinite(out[:B]).all() and float(out[B]) == CANARY
    ref = (a[:, :, :nx].double() ** 2).sum(dim=(1, 2))
    assert torch.allclose(out[:B], ref, rtol=1e-12)
