"""Finite-difference operators and ghost fills on padded 2-D arrays, float64
NumPy restatement (test infrastructure).

Follows moljax/core/operators.py:159-322 (outputs are zero in the ghost ring)
and moljax/core/bc.py:202-307 (x ghosts first, then full ghost rows).
Pinned by the exact eigenvalue (2cos(k h)-2)/h^2 of sine modes and by
second-order convergence on smooth fields (tests/test_oracle_golden.py).
"""
import numpy as np


def d2_central_2d(f, dy, dx, axis):
    """operators.py:271-308."""
    out = np.zeros_like(f)
    if axis == 0:
        out[1:-1, 1:-1] = (f[2:, 1:-1] - 2.0 * f[1:-1, 1:-1] + f[:-2, 1:-1]) / dy ** 2
    else:
        out[1:-1, 1:-1] = (f[1:-1, 2:] - 2.0 * f[1:-1, 1:-1] + f[1:-1, :-2]) / dx ** 2
    return out


def laplacian_2d(f, dy, dx):
    """operators.py:311-322."""
    return d2_central_2d(f, dy, dx, 0) + d2_central_2d(f, dy, dx, 1)


def d1_central_2d(f, dy, dx, axis):
    """operators.py:159-197."""
    out = np.zeros_like(f)
    if axis == 0:
        out[1:-1, 1:-1] = (f[2:, 1:-1] - f[:-2, 1:-1]) / (2.0 * dy)
    else:
        out[1:-1, 1:-1] = (f[1:-1, 2:] - f[1:-1, :-2]) / (2.0 * dx)
    return out


def d1_upwind_2d(f, v, dy, dx, axis):
    """operators.py:200-268: backward where v >= 0 else forward."""
    out = np.zeros_like(f)
    c = f[1:-1, 1:-1]
    if axis == 0:
        bwd = (c - f[:-2, 1:-1]) / dy
        fwd = (f[2:, 1:-1] - c) / dy
    else:
        bwd = (c - f[1:-1, :-2]) / dx
        fwd = (f[1:-1, 2:] - c) / dx
    vi = v if np.ndim(v) == 0 else v[1:-1, 1:-1]
    out[1:-1, 1:-1] = np.where(vi >= 0, bwd, fwd)
    return out


def apply_bc(f, kind, values=(0.0, 0.0, 0.0, 0.0), dy=1.0, dx=1.0):
    """bc.py:202-307 for one field; values = (left, right, bottom, top)."""
    f = f.copy()
    ny, nx = f.shape[0] - 2, f.shape[1] - 2
    vl, vr, vb, vt = values
    if kind == 0:
        f[:, 0] = f[:, nx]
        f[:, nx + 1] = f[:, 1]
        f[0, :] = f[ny, :]
        f[ny + 1, :] = f[1, :]
    elif kind == 1:
        f[:, 0] = 2.0 * vl - f[:, 1]
        f[:, nx + 1] = 2.0 * vr - f[:, nx]
        f[0, :] = 2.0 * vb - f[1, :]
        f[ny + 1, :] = 2.0 * vt - f[ny, :]
    else:
        f[:, 0] = f[:, 1] - 2.0 * dx * vl
        f[:, nx + 1] = f[:, nx] + 2.0 * dx * vr
        f[0, :] = f[1, :] - 2.0 * dy * vb
        f[ny + 1, :] = f[ny, :] + 2.0 * dy * vt
    return f
