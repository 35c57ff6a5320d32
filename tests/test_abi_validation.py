"""Argument checking of the C ABI without a GPU: every case below is
rejected (or is a no-op) before the library touches the device, with the
documented MJX_E_* code and a message in mjx_last_error."""
import ctypes

import pytest

E_ARG, E_SIZE = -1, -2
P = ctypes.c_void_p(64)  # a non-null stand-in pointer; never dereferenced on these paths


@pytest.fixture(scope="module")
def lib():
    from moljax_amd import _hip
    if not _hip.LIB_PATH.exists():
        pytest.fail("libmoljax_hip.so not built: run __graft_entry__.build()")
    return _hip.load_library(require_gpu=False)


def _msg(lib):
    return lib.mjx_last_error().decode()


def test_etd_exp_checks(lib):
    call = lambda order=1, n=64, prev=0, nu=P, steps=3: lib.mjx_gs_etd_exp(
        1, order, 1, n, n, P, P, n, n * n, 0, P, P, prev, nu, nu, steps, P, None)
    assert call(order=3) == E_ARG and "order" in _msg(lib)
    assert call(order=1, prev=1) == E_ARG
    assert call(order=2, prev=3) == E_ARG
    assert call(n=3000) == E_SIZE and "[2, 2048]" in _msg(lib)
    assert call(order=2, prev=1, nu=None) == E_ARG and "nu_prev" in _msg(lib)
    assert call(steps=-1) == E_ARG
    assert call(order=2, prev=1, steps=0) == 0  # nothing to do: no device work
    assert lib.mjx_etd_phi_coeffs(1, 1, 63, P, P, 1e-5, 1e-5, 1.0, P, None) == E_ARG
    assert lib.mjx_etdrk4_coeffs_form(1, 2, 64, 64, P, P, 1e-5, 1e-5, 1.0, P, None) == E_ARG
    assert lib.mjx_etd_phi_coeff_elems(64, 64) == 3 * 64 * 66
    # any-length ETD workspace: six planes of ny * (nx + 2)
    assert lib.mjx_etdrk4_workspace_bytes(0, 2, 48, 50) == 8 * 6 * 2 * 48 * 52
    # power-of-two square ETD workspace: four packed planes of n^2, the last
    # with room for ETD2's separated N^ plane of n (n + 2)
    assert lib.mjx_etdrk4_workspace_bytes(0, 2, 64, 64) == 8 * 2 * (4 * 64 * 64 + 2 * 64)
    assert lib.mjx_etdrk4_workspace_bytes(1, 3, 512, 512) == 16 * 3 * (4 * 512 * 512 + 2 * 512)


def test_imex_and_batch_scheme_checks(lib):
    assert lib.mjx_gs_imex(1, 7, 1, 64, 64, P, P, 66, 66 * 66, 1, P, P, None, P, 1, P, None) == E_ARG
    # the increment form (scheme 4) takes power-of-two sides only
    assert lib.mjx_gs_imex(0, 4, 1, 96, 96, P, P, 98, 98 * 98, 1, P, P, None, P, 1, P, None) == E_SIZE
    assert "power" in _msg(lib)
    plan = ctypes.c_void_p()
    n = (ctypes.c_int * 1)(64)
    assert lib.mjx_batch_create(1, 5, 1, n, n, None, None, None, None, None, None, None, 4,
                                ctypes.byref(plan)) == E_ARG
    assert "scheme" in _msg(lib) and plan.value is None


def test_r2r_checks(lib):
    sv = (ctypes.c_double * 4)(0.1, 0.2, 0.3, 0.4)
    bc2 = lambda kind=2, n=64, mmode=0, side=sv, dx=0.1: lib.mjx_r2r_solve2d_bc(
        1, kind, 1, n, n, P, n, n * n, P, n, n * n, P, P, P, mmode, side, dx, dx, P, None)
    assert bc2(kind=0) == E_ARG and "kind 1" in _msg(lib)
    assert bc2(kind=3) == E_ARG
    assert bc2(mmode=2) == E_ARG
    assert bc2(n=3000) == E_SIZE and "[2, 2048]" in _msg(lib)
    assert bc2(n=1) == E_SIZE
    assert bc2(side=None) == E_ARG
    assert bc2(dx=0.0) == E_ARG
    assert lib.mjx_r2r_solve1d_bc(1, 2, 1, 4097, P, 4097, P, 4097, P, P, 0, sv, 0.1, None) == E_SIZE
    assert lib.mjx_r2r_transform1d(1, 3, 0, 1, 64, P, 64, P, 64, 1.0, None) == E_ARG
    assert lib.mjx_r2r_solve2d(1, 2, 1, 2049, 64, P, 64, 64 * 2049, P, 64, 64 * 2049, P, P, P, 0, P, None) == E_SIZE
    assert lib.mjx_r2r_transform1d(1, 0, 0, 1, 2049, P, 2049, P, 2049, 1.0, None) == E_SIZE


def test_benchmark_method_checks(lib):
    rk = lambda n=64, inst=P, steps=2: lib.mjx_gs_rk4_spectral(1, 1, n, n, P, P, n, n * n, P, P, inst, steps, P,
                                                                None)
    assert rk(n=48) == E_SIZE and "powers of two" in _msg(lib)
    assert rk(inst=None) == E_ARG
    assert rk(steps=-2) == E_ARG
    assert rk(steps=0) == 0
    cn = lambda max_iter=10, lap=P: lib.mjx_gs_cn_fixed_point(1, 1, 64, 64, P, P, 64, 64 * 64, P, P, lap, P, 1e-8,
                                                               max_iter, 2, None, P, None)
    assert cn(max_iter=0) == E_ARG and "max_iter" in _msg(lib)
    assert cn(lap=None) == E_ARG
    assert lib.mjx_gs_bench_workspace_bytes(1, 2, 64, 32) == 6 * 8 * 2 * 64 * 32 + 16 * 2 * 64 + 4 * 2
    assert lib.mjx_reduce2d(1, 9, 1, 4, 4, P, 4, 16, None, 4, 16, 0.0, 0.0, P, 0, P, None) == E_ARG
    assert lib.mjx_reduce_finish(0, 0, P, 1, P, None) == E_ARG
    assert lib.mjx_reduce_finish(9, 4, P, 1, P, None) == E_ARG
    assert lib.mjx_reduce_finish(0, 4, None, 1, P, None) == E_ARG


def test_spectral_any_length_checks(lib):
    # power-of-two grids keep the packed layout, other sizes the half-spectra
    assert lib.mjx_spectral_workspace_bytes(0, 2, 64, 32) == 2 * 64 * 32 * 8
    assert lib.mjx_spectral_workspace_bytes(0, 2, 50, 30) == 2 * 50 * 32 * 8
    assert lib.mjx_spectral_workspace_bytes(1, 1, 64, 33) == 64 * 34 * 16
    pair = lambda ny, nx: lib.mjx_spectral_pair(1, 1, ny, nx, P, None, nx, nx * ny, P, None, nx, nx * ny, 0, P, P,
                                                None, P, 1, P, None)
    assert pair(3000, 32) == E_SIZE and "[2, 2048]" in _msg(lib)
    assert pair(32, 1) == E_SIZE
    assert lib.mjx_gs_imex_strang(1, 1, 2049, 2049, P, P, 2049, 2049 * 2049, 0, P, P, None, P, 1, P, None) == E_SIZE


def test_fd_op_checks(lib):
    # an upwind velocity array must cover the padded row and must not overlap between instances
    fd = lambda B=2, vel=P, vp=66, vs=66 * 66, nx=64: lib.mjx_fd_op2d(
        1, 3, B, 64, nx, P, 66, 66 * 66, P, 66, 66 * 66, 1, 0.1, 0.1, vel, vp, vs, 0.0, None)
    assert fd(vp=65) == E_ARG and "velocity pitch" in _msg(lib)
    assert fd(vs=66) == E_ARG
    assert lib.mjx_fd_op2d(1, 3, 1, 4, 2 ** 31 - 66, P, 2 ** 31 - 1, 0, P, 2 ** 31 - 1, 0, 1, 0.1, 0.1,
                           None, 0, 0, 1.0, None) == -4 and "too large" in _msg(lib)

