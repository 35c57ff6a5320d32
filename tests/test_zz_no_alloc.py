"""mjx_gs_imex_strang_noalloc (the allocation-free Gray-Scott IMEX-Strang
entry point, workspace in a fixed device array) against mjx_gs_imex_strang
with a caller-allocated workspace, through the C ABI: bitwise equal results
on two calls in a row, the second with a smaller input (the scratch keeps
the first call's contents), in FP32 and FP64, and a third call on a
mixed-radix side (250); oversize requests and sides without a Stockham plan
are refused with MJX_E_LIMIT / MJX_E_SIZE before any launch."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _state(n, B, dtype, seed0):
    from moljax_amd.workloads import gray_scott_batch
    return gray_scott_batch(n, n, list(range(seed0, seed0 + B)), dtype=dtype)


def _args(U, V, n, dtype):
    from moljax_amd import _hip
    from moljax_amd.batch import fd_symbol_1d
    from moljax_amd.core._arrays import inst_table
    lam = torch.from_numpy(fd_symbol_1d(n, 2.5)).to("cuda", dtype)
    B = U.shape[0]
    rows = inst_table([[0.5 * 2e-5, 0.5 * 1e-5, 0, 0, 0.035 + 0.001 * b, 0.065, 1.0] for b in range(B)])
    W = n + 2
    return lam, rows, (_hip.dtype_code(dtype), B, n, n, _hip.ptr(U[:, 1:, 1:]), _hip.ptr(V[:, 1:, 1:]), W, W * W, 1,
                       _hip.ptr(lam), _hip.ptr(lam), None, _hip.ptr(rows))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_noalloc_equals_allocating_entry_point(gpu, dtype):
    from moljax_amd import _hip
    from moljax_amd.core._arrays import spectral_work
    lib = _hip.lib()
    assert lib.mjx_init() == 0
    st = _hip.stream_ptr()
    # second call smaller than the first: the scratch holds stale data from
    # the first call when the second starts
    for n, B, steps, seed0 in ((256, 16, 7, 0), (64, 3, 5, 40), (250, 4, 5, 80)):
        print(f"noalloc {dtype} n={n} B={B} steps={steps}", flush=True)
        Ua, Va = _state(n, B, dtype, seed0)
        Ub, Vb = Ua.clone(), Va.clone()
        lam_a, rows_a, args_a = _args(Ua, Va, n, dtype)
        work = spectral_work(dtype, B, n, n)
        rc = lib.mjx_gs_imex_strang(*args_a, steps, _hip.ptr(work), st)
        _hip.check(rc, "mjx_gs_imex_strang")
        lam_b, rows_b, args_b = _args(Ub, Vb, n, dtype)
        rc = lib.mjx_gs_imex_strang_noalloc(*args_b, steps, st)
        _hip.check(rc, "mjx_gs_imex_strang_noalloc")
        torch.cuda.synchronize()
        assert torch.isfinite(Ua).all() and torch.isfinite(Va).all()
        assert torch.equal(Ua, Ub) and torch.equal(Va, Vb), f"n={n} B={B}: no-alloc result differs"


def test_noalloc_refuses_oversize_and_other_sides(gpu):
    from moljax_amd import _hip
    lib = _hip.lib()
    assert lib.mjx_init() == 0
    P = 64  # a non-null stand-in pointer: these calls return before any launch
    limit = 64 << 20
    # FP32 256^2 needs 512 KiB per instance: 129 instances exceed 64 MiB
    assert lib.mjx_spectral_workspace_bytes(0, 129, 256, 256) > limit
    assert lib.mjx_spectral_workspace_bytes(0, 128, 256, 256) <= limit
    rc = lib.mjx_gs_imex_strang_noalloc(0, 129, 256, 256, P, P, 258, 258 * 258, 1, P, P, None, P, 3,
                                        _hip.stream_ptr())
    assert rc == _hip.MJX_E_LIMIT and "scratch" in lib.mjx_last_error().decode()
    # 252 has no Stockham plan (chirp-z sides are not served)
    rc = lib.mjx_gs_imex_strang_noalloc(0, 2, 252, 252, P, P, 254, 254 * 254, 1, P, P, None, P, 3,
                                        _hip.stream_ptr())
    assert rc == _hip.MJX_E_SIZE
