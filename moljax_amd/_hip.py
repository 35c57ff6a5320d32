"""ctypes bridge to libmoljax_hip.so (the C ABI declared in include/moljax_hip.h).

torch is imported before the library is opened so that the HIP runtime torch
loaded (soname libamdhip64.so.7) is the one the library binds to: device
pointers and streams from torch are then valid in the kernels.  There is no
fallback: every entry point raises if the library or a GPU is missing.
"""
import ctypes
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "libmoljax_hip.so"

MJX_F32 = 0
MJX_F64 = 1
MJX_E_ARG = -1
MJX_E_SIZE = -2
MJX_E_HIP = -3
MJX_E_LIMIT = -4

_i = ctypes.c_int
_ll = ctypes.c_longlong
_p = ctypes.c_void_p
_d = ctypes.c_double
_dp = ctypes.POINTER(ctypes.c_double)  # host array

# name -> (restype, argtypes); must match include/moljax_hip.h
PROTOTYPES = {
    "mjx_version": (_i, []),
    "mjx_last_error": (ctypes.c_char_p, []),
    "mjx_init": (_i, []),
    "mjx_spectral_workspace_bytes": (_ll, [_i, _i, _i, _i]),
    "mjx_spectral_pair": (_i, [_i, _i, _i, _i, _p, _p, _ll, _ll, _p, _p, _ll, _ll, _i, _p, _p, _p, _p, _i, _p, _p]),
    "mjx_gs_imex_strang": (_i, [_i, _i, _i, _i, _p, _p, _ll, _ll, _i, _p, _p, _p, _p, _i, _p, _p]),
    "mjx_gs_imex_strang_noalloc": (_i, [_i, _i, _i, _i, _p, _p, _ll, _ll, _i, _p, _p, _p, _p, _i, _p]),
    "mjx_gs_imex_strang_var": (_i, [_i, _i, _i, _i, _p, _p, _ll, _ll, _i, _p, _p, _p, _p, _i, _p, _p]),
    "mjx_gs_imex": (_i, [_i, _i, _i, _i, _i, _p, _p, _ll, _ll, _i, _p, _p, _p, _p, _i, _p, _p]),
    "mjx_gs_adaptive_workspace_bytes": (_ll, [_i, _i, _i, _i, _i]),
    "mjx_gs_explicit_workspace_bytes": (_ll, [_i, _i, _i, _i]),
    "mjx_gs_explicit": (_i, [_i, _i, _i, _i, _i, _p, _p, _ll, _ll, _i, _d, _d, _p, _i, _p, _p]),
    "mjx_gs_imex_adaptive": (_i, [_i, _i, _i, _i, _i, _p, _p, _ll, _ll, _i, _p, _p, _p, _p, _dp, _p, _p, _p, _p,
                                  _i, _ll, ctypes.POINTER(ctypes.c_longlong), _p, _p]),
    "mjx_reduce_workspace_elems": (_ll, [_i]),
    "mjx_reduce2d": (_i, [_i, _i, _i, _i, _i, _p, _ll, _ll, _p, _ll, _ll, _d, _d, _p, _i, _p, _p]),
    "mjx_reduce_finish": (_i, [_i, _i, _p, _i, _p, _p]),
    "mjx_batch_create": (_i, [_i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _p]),
    "mjx_batch_create_mixed": (_i, [_i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _p]),
    "mjx_batch_run": (_i, [_p, _i, _i, _p]),
    "mjx_batch_destroy": (_i, [_p]),
    "mjx_autotune": (_i, [_i, _i, _i, _i, _i, ctypes.POINTER(ctypes.c_int), _p]),
    "mjx_tune_get": (_i, [_i, _i, _i]),
    "mjx_tune_set": (_i, [_i, _i, _i, _i]),
    "mjx_fd_op2d": (_i, [_i, _i, _i, _i, _i, _p, _ll, _ll, _p, _ll, _ll, _i, _d, _d, _p, _ll, _ll, _d, _p]),
    "mjx_bc_fill2d": (_i, [_i, _i, _i, _i, _p, _ll, _ll, _i, _dp, _d, _d, _p]),
    "mjx_gs_reaction": (_i, [_i, _ll, _p, _p, _p, _p, _d, _d, _p]),
    "mjx_r2r_solve2d": (_i, [_i, _i, _i, _i, _i, _p, _ll, _ll, _p, _ll, _ll, _p, _p, _p, _i, _p, _p]),
    "mjx_r2r_solve2d_bc": (_i, [_i, _i, _i, _i, _i, _p, _ll, _ll, _p, _ll, _ll, _p, _p, _p, _i, _p, _d, _d, _p, _p]),
    "mjx_r2r_solve1d_bc": (_i, [_i, _i, _i, _i, _p, _ll, _p, _ll, _p, _p, _i, _p, _d, _p]),
    "mjx_r2r_solve1d": (_i, [_i, _i, _i, _i, _p, _ll, _p, _ll, _p, _p, _i, _p]),
    "mjx_axpby": (_i, [_i, _ll, _d, _p, _d, _p, _d, _p, _p]),
    "mjx_jacobi2d": (_i, [_i, _i, _i, _i, _p, _ll, _ll, _p, _ll, _ll, _p, _d, _d, _d, _d, _i, _p]),
    "mjx_r2r_etd1": (_i, [_i, _i, _i, _i, _p, _p, _ll, _p, _ll, _p, _d, _p]),
    "mjx_r2r_transform1d": (_i, [_i, _i, _i, _i, _i, _p, _ll, _p, _ll, _d, _p]),
    "mjx_etdrk4_coeff_elems": (_ll, [_i, _i]),
    "mjx_etdrk4_coeffs": (_i, [_i, _i, _i, _p, _p, _d, _d, _d, _p, _p]),
    "mjx_etdrk4_coeffs_form": (_i, [_i, _i, _i, _i, _p, _p, _d, _d, _d, _p, _p]),
    "mjx_etdrk4_workspace_bytes": (_ll, [_i, _i, _i, _i]),
    "mjx_gs_etdrk4": (_i, [_i, _i, _i, _i, _p, _p, _ll, _ll, _i, _p, _p, _i, _p, _p]),
    "mjx_gs_bench_workspace_bytes": (_ll, [_i, _i, _i, _i]),
    "mjx_gs_rk4_spectral": (_i, [_i, _i, _i, _i, _p, _p, _ll, _ll, _p, _p, _p, _i, _p, _p]),
    "mjx_gs_cn_fixed_point": (_i, [_i, _i, _i, _i, _p, _p, _ll, _ll, _p, _p, _p, _p, _d, _i, _i, _p, _p, _p]),
    "mjx_phi_tables": (_i, [_i, _i, _p, _d, _i, _i, _p, _p, _p, _p, _p]),
    "mjx_etd_phi_coeff_elems": (_ll, [_i, _i]),
    "mjx_etd_phi_coeffs": (_i, [_i, _i, _i, _p, _p, _d, _d, _d, _p, _p]),
    "mjx_gs_etd_exp": (_i, [_i, _i, _i, _i, _i, _p, _p, _ll, _ll, _i, _p, _p, _i, _p, _p, _i, _p, _p]),
    "mjx_nk_workspace_bytes": (_ll, [_i, _i, _i]),
    "mjx_rd_newton_krylov": (_i, [_i, _i, _i, _d, _d, _d, _d, _d, _dp, _p, _p, _p, _p, _dp,
                                  ctypes.POINTER(ctypes.c_int), _dp, _p, _p]),
    "mjx_rd_residual": (_i, [_i, _i, _i, _d, _d, _d, _d, _d, _d, _dp, _p, _p, _p, _p, _p]),
    "mjx_rd_jvp": (_i, [_i, _i, _i, _i, _d, _d, _d, _d, _d, _d, _dp, _p, _p, _p, _p]),
    "mjx_dot_partial_elems": (_ll, [_i]),
    "mjx_dot_f64": (_i, [_ll, _i, _p, _p, _p, _p, _p]),
    "mjx_kernel_timing_enable": (_i, [_i]),
    "mjx_kernel_timing": (_i, [_i, _dp, ctypes.POINTER(ctypes.c_longlong)]),
    "mjx_kernel_timing_samples": (_ll, [_i, _dp, _ll]),
}

_lib = None
_gpu_checked = False


class MjxError(RuntimeError):
    """A kernel-library call failed (HIP runtime error or capacity limit)."""


def load_library(require_gpu: bool = True):
    """Open the library (once) and bind prototypes.

    With ``require_gpu`` (the default for every compute path) a missing GPU is
    an error; tests that only inspect exported symbols pass False.
    """
    global _lib, _gpu_checked
    if require_gpu and not _gpu_checked:
        if not torch.cuda.is_available():
            raise MjxError("moljax_amd needs a ROCm GPU; torch.cuda.is_available() is False")
        _gpu_checked = True  # a process keeps its GPUs: checked once
    if _lib is None:
        if not LIB_PATH.exists():
            raise MjxError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def lib():
    return load_library(True)


def last_error() -> str:
    msg = lib().mjx_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str):
    if rc == 0:
        return
    msg = last_error()
    if rc in (MJX_E_ARG, MJX_E_SIZE):
        raise ValueError(f"{what}: {msg}")
    raise MjxError(f"{what} failed ({rc}): {msg}")


def dtype_code(dt: torch.dtype) -> int:
    if dt == torch.float32:
        return MJX_F32
    if dt == torch.float64:
        return MJX_F64
    raise ValueError(f"unsupported dtype {dt}: the kernels take float32 or float64")


def stream_ptr(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t) -> int:
    return 0 if t is None else t.data_ptr()


def kernel_timing_samples(kind: int):
    """Per-launch durations (ms) of a hot kernel since timing was enabled."""
    n = lib().mjx_kernel_timing_samples(kind, None, 0)
    if n < 0:
        check(int(n), "mjx_kernel_timing_samples")
    buf = (ctypes.c_double * max(int(n), 1))()
    lib().mjx_kernel_timing_samples(kind, buf, n)
    return list(buf[:n])


def kernel_timing(kind: int):
    """(total ms, launch count) of a hot kernel since timing was enabled."""
    ms = ctypes.c_double(0.0)
    n = ctypes.c_longlong(0)
    check(lib().mjx_kernel_timing(kind, ctypes.byref(ms), ctypes.byref(n)), "mjx_kernel_timing")
    return ms.value, n.value

