"""FP64-accumulated reductions (tree_norm2, tree_vdot, tree_norm_inf,
scaled_error_norm, is_finite) on the GPU, pinned by math.fsum brute force on
the same values (moljax/core/state.py:115-189,312-343)."""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


class _G:
    """Minimal 2-D grid: one ghost ring."""
    def __init__(self, ny, nx):
        self.ny, self.nx = ny, nx
        self.interior_slice = (slice(1, ny + 1), slice(1, nx + 1))


def _state(rng, ny, nx, dtype):
    return {k: torch.from_numpy(rng.standard_normal((ny + 2, nx + 2)) * s).to("cuda", dtype)
            for k, s in (("u", 1.0), ("v", 1e-3))}


def _np(t):
    return t.detach().cpu().double().numpy()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("shape", [(1, 7), (33, 65), (256, 256)])
def test_norms_match_fsum(gpu, dtype, shape):
    from moljax_amd.core.state import scaled_error_norm, tree_norm2, tree_norm_inf, tree_vdot
    ny, nx = shape
    rng = np.random.default_rng(ny * 1000 + nx)
    g = _G(ny, nx)
    a, b = _state(rng, ny, nx, dtype), _state(rng, ny, nx, dtype)
    ia = {k: _np(v)[1:-1, 1:-1] for k, v in a.items()}
    ib = {k: _np(v)[1:-1, 1:-1] for k, v in b.items()}
    n2 = math.sqrt(math.fsum(np.concatenate([(x * x).ravel() for x in ia.values()])))
    assert float(tree_norm2(a, g)) == pytest.approx(n2, rel=1e-14)
    dot = math.fsum(np.concatenate([(ia[k] * ib[k]).ravel() for k in ia]))
    assert abs(float(tree_vdot(a, b, g)) - dot) <= 1e-14 * math.fsum(
        np.concatenate([np.abs(ia[k] * ib[k]).ravel() for k in ia]))
    assert float(tree_norm_inf(a, g)) == max(np.max(np.abs(x)) for x in ia.values())
    full = math.sqrt(math.fsum(np.concatenate([(_np(x) ** 2).ravel() for x in a.values()])))
    assert float(tree_norm2(a, g, interior_only=False)) == pytest.approx(full, rel=1e-14)
    atol, rtol = 1e-6, 1e-4
    terms = np.concatenate([((ia[k] / (atol + rtol * np.abs(ib[k]))) ** 2).ravel() for k in ia])
    sen = math.sqrt(math.fsum(terms) / terms.size)
    assert float(scaled_error_norm(a, b, g, atol, rtol)) == pytest.approx(sen, rel=1e-13)


def test_bitwise_reproducible_and_nonfinite(gpu):
    from moljax_amd.core.state import is_finite, tree_norm2, tree_norm_inf
    rng = np.random.default_rng(7)
    g = _G(300, 200)
    a = _state(rng, 300, 200, torch.float64)
    r = [float(tree_norm2(a, g)) for _ in range(3)]
    assert r[0] == r[1] == r[2]
    assert is_finite(a)
    a["v"][5, 7] = float("nan")
    assert not is_finite(a)
    assert math.isnan(float(tree_norm_inf(a, g)))
    a["v"][5, 7] = float("inf")
    assert not is_finite(a)


def test_batched_reduction_per_instance(gpu):
    from moljax_amd.core.state import RED_SUMSQ, _reduce
    rng = np.random.default_rng(3)
    x = torch.from_numpy(rng.standard_normal((5, 34, 66))).cuda()
    inner = x[:, 1:-1, 1:-1]
    out = _reduce(RED_SUMSQ, [inner]).cpu().numpy()
    for b in range(5):
        assert out[b] == pytest.approx(math.fsum((inner[b].cpu().numpy() ** 2).ravel()), rel=1e-14)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_interior_sum_mean_max_and_history(gpu, dtype):
    """utils.py:63-153,222-262: interior sum / mean / max |x| (native FP64
    reductions) against math.fsum, and the history buffers."""
    from moljax_amd.core.utils import (allocate_scalar_history, allocate_state_history, create_padded_array,
                                       max_abs_interior, mean_interior, save_to_history, set_interior,
                                       sum_interior)
    rng = np.random.default_rng(17)
    g = _G(48, 80)
    f = torch.from_numpy(rng.standard_normal((50, 82))).to("cuda", dtype)
    x = _np(f)[1:-1, 1:-1]
    assert sum_interior(f, g) == pytest.approx(math.fsum(x.ravel()), rel=1e-12, abs=1e-12)
    assert mean_interior(f, g) == pytest.approx(math.fsum(x.ravel()) / x.size, rel=1e-12, abs=1e-14)
    assert max_abs_interior(f, g) == float(np.abs(x).max())
    batch = torch.stack([f, 2 * f])
    assert max_abs_interior(batch, g) == 2 * float(np.abs(x).max())
    assert sum_interior(batch, g) == pytest.approx(3 * math.fsum(x.ravel()), rel=1e-12, abs=1e-12)
    g.ny_total, g.nx_total = 50, 82
    p = create_padded_array(g, 0.5, dtype)
    assert p.shape == (50, 82) and float(p.sum()) == 0.5 * 50 * 82
    q = set_interior(p, torch.from_numpy(x).to("cuda", dtype), g)
    assert torch.equal(q[1:-1, 1:-1], f[1:-1, 1:-1]) and float(q[0, 0]) == 0.5 and float(p[1, 1]) == 0.5
    hist = allocate_state_history({"u": f, "v": f}, g, 4, dtype=dtype)
    assert hist["u"].shape == (4, 48, 80) and allocate_scalar_history(4).shape == (4,)
    save_to_history(hist, {"u": f, "v": 3 * f}, g, 2)
    assert torch.equal(hist["u"][2], f[1:-1, 1:-1]) and torch.equal(hist["v"][2], 3 * f[1:-1, 1:-1])
    assert float(hist["u"][1].abs().max()) == 0.0


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_tree_algebra_and_neumann_compatibility(gpu, dtype):
    """state.py:27-81 (tree_add / sub / scale / axpy on the native a x + b y + c
    kernel) and fft_nonperiodic.py:440-461 (compatibility check and projection
    with the native sum)."""
    from moljax_amd.core.fft_nonperiodic import check_compatibility_neumann, project_to_compatible
    from moljax_amd.core.state import tree_add, tree_axpy, tree_scale, tree_sub
    rng = np.random.default_rng(5)
    a = {"u": rng.standard_normal((10, 12)), "v": rng.standard_normal((3, 10, 12))}
    b = {"u": rng.standard_normal((10, 12)), "v": rng.standard_normal((3, 10, 12))}
    ta = {k: torch.from_numpy(v).to("cuda", dtype) for k, v in a.items()}
    tb = {k: torch.from_numpy(v).to("cuda", dtype) for k, v in b.items()}
    ra = {k: _np(v) for k, v in ta.items()}
    rb = {k: _np(v) for k, v in tb.items()}
    tol = 4e-16 if dtype == torch.float64 else 1e-6  # fused multiply-adds round once
    for got, want in ((tree_add(ta, tb), {k: ra[k] + rb[k] for k in ra}),
                      (tree_sub(ta, tb), {k: ra[k] - rb[k] for k in ra}),
                      (tree_scale(ta, 0.3), {k: 0.3 * ra[k] for k in ra}),
                      (tree_axpy(ta, -1.7, tb), {k: ra[k] + (-1.7) * rb[k] for k in ra})):
        for k in want:
            assert got[k].shape == ta[k].shape and got[k].dtype == dtype
            assert np.max(np.abs(_np(got[k]) - want[k])) <= tol * (1 + np.max(np.abs(want[k]))), k
    assert torch.equal(ta["u"], torch.from_numpy(a["u"]).to("cuda", dtype))  # inputs untouched
    r = torch.from_numpy(rng.standard_normal(64)).to("cuda", dtype)
    p = project_to_compatible(r)
    assert check_compatibility_neumann(p, tol=1e-10 if dtype == torch.float64 else 1e-4)
    assert not check_compatibility_neumann(r + 1.0)
    assert np.allclose(_np(p), _np(r) - math.fsum(_np(r)) / 64, atol=1e-14 if dtype == torch.float64 else 1e-6)
