"""Seeded synthetic Gray-Scott workloads (initial conditions and batches).

IC on interior cells (ny, nx): u = 1, v = 0, a centred square of ``square``
cells with u = 0.5, v = 0.25, plus uniform noise in [-noise, noise) on u then
v drawn from numpy.random.default_rng(seed); the domain and square convention
of examples/gray_scott_2d.py:42-73 with a reproducible noise source.
Arrays are returned ghost-padded with the periodic ring filled.
"""
from typing import Sequence, Tuple

import numpy as np
import torch


def _wrap(a: np.ndarray) -> np.ndarray:
    p = np.zeros((a.shape[0] + 2, a.shape[1] + 2), dtype=a.dtype)
    p[1:-1, 1:-1] = a
    p[:, 0] = p[:, -2]
    p[:, -1] = p[:, 1]
    p[0, :] = p[-2, :]
    p[-1, :] = p[1, :]
    return p


def gray_scott_ic(ny: int, nx: int, seed: int, square: int = 20, noise: float = 0.01) -> Tuple[np.ndarray, np.ndarray]:
    rng = np.random.default_rng(seed)
    u = np.ones((ny, nx))
    v = np.zeros((ny, nx))
    cy, cx, h = ny // 2, nx // 2, square // 2
    u[cy - h:cy - h + square, cx - h:cx - h + square] = 0.5
    v[cy - h:cy - h + square, cx - h:cx - h + square] = 0.25
    u = u + rng.uniform(-noise, noise, size=(ny, nx))
    v = v + rng.uniform(-noise, noise, size=(ny, nx))
    return _wrap(u), _wrap(v)


def gray_scott_batch(ny: int, nx: int, seeds: Sequence[int], dtype=torch.float32, device="cuda",
                     square: int = 20) -> Tuple[torch.Tensor, torch.Tensor]:
    """(B, ny+2, nx+2) padded u, v on the device for the given seeds."""
    B = len(seeds)
    U = torch.empty((B, ny + 2, nx + 2), dtype=dtype, device=device)
    V = torch.empty_like(U)
    for i, s in enumerate(seeds):
        u, v = gray_scott_ic(ny, nx, int(s), square)
        U[i].copy_(torch.from_numpy(u))
        V[i].copy_(torch.from_numpy(v))
    return U, V
