"""FD operators, ghost fill and reaction kernels vs the float64 oracle.
The formulas are exact, so float64 results agree to the last few ulps."""
import numpy as np
import pytest
import torch

from oracle import fd as ofd

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().double().cpu().numpy()


def _field(ny, nx, seed=0):
    return np.random.default_rng(seed).standard_normal((ny + 2, nx + 2))


@pytest.mark.parametrize("ny,nx", [(16, 16), (33, 70), (256, 256), (24, 37), (70, 129)])
def test_fd_operators(gpu, ny, nx):
    from moljax_amd.core import Grid2D
    from moljax_amd.core import operators as op
    grid = Grid2D.uniform(nx, ny, 0.0, 1.3, 0.0, 0.7)
    f = _field(ny, nx, ny)
    t = torch.from_numpy(f).cuda()
    dy, dx = grid.dy, grid.dx
    assert np.allclose(_np(op.laplacian_2d(t, grid)), ofd.laplacian_2d(f, dy, dx), rtol=1e-13, atol=1e-9)
    for axis in (0, 1):
        assert np.allclose(_np(op.d2_central_2d(t, grid, axis)), ofd.d2_central_2d(f, dy, dx, axis),
                           rtol=1e-13, atol=1e-9)
        assert np.allclose(_np(op.d1_central_2d(t, grid, axis)), ofd.d1_central_2d(f, dy, dx, axis),
                           rtol=1e-13, atol=1e-11)
        vel = _field(ny, nx, 99)
        assert np.allclose(_np(op.d1_upwind_2d(t, torch.from_numpy(vel).cuda(), grid, axis)),
                           ofd.d1_upwind_2d(f, vel, dy, dx, axis), rtol=1e-13, atol=1e-11)
        assert np.allclose(_np(op.d1_upwind_2d(t, -0.5, grid, axis)), ofd.d1_upwind_2d(f, -0.5, dy, dx, axis),
                           rtol=1e-13, atol=1e-11)
    out = _np(op.laplacian_2d(t, grid))
    assert np.all(out[0] == 0) and np.all(out[-1] == 0) and np.all(out[:, 0] == 0) and np.all(out[:, -1] == 0)


def test_fd_batched_and_float32(gpu):
    from moljax_amd.core import Grid2D
    from moljax_amd.core import operators as op
    grid = Grid2D.uniform(64, 32, 0.0, 1.0, 0.0, 1.0)
    f = np.stack([_field(32, 64, s) for s in range(3)])
    lap = _np(op.laplacian_2d(torch.from_numpy(f).to("cuda", torch.float32), grid))
    for b in range(3):
        ref = ofd.laplacian_2d(f[b].astype(np.float32).astype(np.float64), grid.dy, grid.dx)
        assert np.allclose(lap[b], ref, rtol=1e-5, atol=1e-2)


@pytest.mark.parametrize("kind,vals", [(0, (0, 0, 0, 0)), (1, (0.3, -1.0, 2.0, 0.5)), (2, (1.0, 2.0, -3.0, 0.25))])
def test_bc_fill(gpu, kind, vals):
    from moljax_amd.core import Grid2D, FieldBCSpec, BCType
    from moljax_amd.core.bc_fill import apply_bc
    grid = Grid2D.uniform(40, 24, 0.0, 2.0, 0.0, 1.0)
    f = _field(24, 40, 5)
    names = ["left", "right", "bottom", "top"]
    if kind == 1:
        spec = FieldBCSpec.dirichlet(lambda X, Y, t, p: dict(zip(names, vals)))
    elif kind == 2:
        spec = FieldBCSpec.neumann(lambda X, Y, t, p: dict(zip(names, vals)))
    else:
        spec = FieldBCSpec.periodic()
    out = apply_bc({"u": torch.from_numpy(f).cuda()}, grid, {"u": spec})["u"]
    ref = ofd.apply_bc(f, kind, vals, grid.dy, grid.dx)
    assert np.allclose(_np(out), ref, rtol=1e-14, atol=1e-14)


def test_gs_reaction(gpu):
    from moljax_amd.core.operators import gray_scott_reaction
    rng = np.random.default_rng(3)
    u, v = rng.random((50, 70)), rng.random((50, 70))
    ru, rv = gray_scott_reaction(torch.from_numpy(u).cuda(), torch.from_numpy(v).cuda(), 0.035, 0.065)
    uvv = u * v * v
    assert np.allclose(_np(ru), -uvv + 0.035 * (1 - u), rtol=1e-14, atol=1e-16)
    assert np.allclose(_np(rv), uvv - (0.035 + 0.065) * v, rtol=1e-14, atol=1e-16)


def test_fd_upwind_velocity_broadcast(gpu):
    """A (1, H, W) velocity with a (3, H, W) field applies to every instance
    (checked per instance against the oracle); a (2, H, W) velocity cannot be
    broadcast and raises before any launch."""
    from moljax_amd.core import Grid2D
    from moljax_amd.core import operators as op
    grid = Grid2D.uniform(40, 24, 0.0, 2.0, 0.0, 1.0)
    f = np.stack([_field(24, 40, s) for s in range(3)])
    vel = _field(24, 40, 77)[None]
    for axis in (0, 1):
        out = _np(op.d1_upwind_2d(torch.from_numpy(f).cuda(), torch.from_numpy(vel).cuda(), grid, axis))
        for b in range(3):
            ref = ofd.d1_upwind_2d(f[b], vel[0], grid.dy, grid.dx, axis)
            assert np.allclose(out[b], ref, rtol=1e-13, atol=1e-11)
    with pytest.raises(ValueError, match="broadcast"):
        op.d1_upwind_2d(torch.from_numpy(f).cuda(), torch.from_numpy(np.zeros((2, 26, 42))).cuda(), grid, 0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_fd_strided_views_odd_width(gpu, dtype):
    """Padded fields as views of wider rows (even pitch, odd padded width:
    two-element accesses with a single last column) and contiguous odd
    widths (one-element accesses) give the oracle's values, with the output
    ghost ring zero."""
    from moljax_amd.core import Grid2D
    from moljax_amd.core import operators as op
    ny, nx = 41, 63  # padded width 65
    grid = Grid2D.uniform(nx, ny, 0.0, 1.1, 0.0, 0.9)
    f = np.stack([_field(ny, nx, s) for s in range(3)])
    wide = torch.zeros((3, ny + 2, nx + 3), dtype=dtype, device="cuda")
    wide[:, :, :nx + 2] = torch.from_numpy(f).to("cuda", dtype)
    view = wide[:, :, :nx + 2]
    tol = dict(rtol=1e-13, atol=1e-9) if dtype == torch.float64 else dict(rtol=2e-5, atol=2e-2)
    for t in (view, view.contiguous()):
        out = _np(op.laplacian_2d(t, grid))
        for b in range(3):
            ref = ofd.laplacian_2d(f[b].astype(np.float32).astype(np.float64) if dtype == torch.float32 else f[b],
                                   grid.dy, grid.dx)
            assert np.allclose(out[b], ref, **tol)
            assert np.all(out[b, 0] == 0) and np.all(out[b, -1] == 0)
            assert np.all(out[b, :, 0] == 0) and np.all(out[b, :, -1] == 0)
        d1 = _np(op.d1_central_2d(t, grid, 1))
        for b in range(3):
            ref = ofd.d1_central_2d(f[b].astype(np.float32).astype(np.float64) if dtype == torch.float32 else f[b],
                                    grid.dy, grid.dx, 1)
            assert np.allclose(d1[b], ref, **tol)
