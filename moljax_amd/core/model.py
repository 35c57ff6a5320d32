"""Method-of-lines models (moljax/core/model.py).

``MOLModel`` keeps the reference's structure (grid, BC spec, params, linear
and nonlinear operator tuples, metadata).  Models built by the factories here
carry ``metadata['mjx_kernel']`` naming the fused MI355X kernels that serve
them; the stepping drivers dispatch on it and reject models they have no
kernel for (no generic Python-callable fallback).
"""
from dataclasses import dataclass, field
from typing import Any, Callable, Dict, List, Optional, Tuple

import torch

from .bc import BCSpec, BCType, FieldBCSpec
from .grid import Grid1D, Grid2D, GridType, default_device
from .operators import LinearOp, NonlinearOp, gray_scott_reaction_op, laplacian_2d
from .state import StateDict, tree_add, tree_zeros_like


@dataclass(frozen=True)
class MOLModel:
    """moljax/core/model.py:34-206."""
    grid: GridType
    bc_spec: BCSpec
    params: Dict[str, Any]
    linear_ops: Tuple[LinearOp, ...] = field(default_factory=tuple)
    nonlinear_ops: Tuple[NonlinearOp, ...] = field(default_factory=tuple)
    metadata: Dict[str, Any] = field(default_factory=dict)

    @property
    def dtype(self) -> torch.dtype:
        return self.params.get("dtype", torch.float64)

    @property
    def field_names(self) -> List[str]:
        return list(self.bc_spec.keys())

    def apply_bcs(self, state: StateDict, t: float = 0.0) -> StateDict:
        from .bc_fill import apply_bc
        return apply_bc(state, self.grid, self.bc_spec, t, self.params)

    def linear_rhs(self, state: StateDict, t: float) -> StateDict:
        if len(self.linear_ops) == 0:
            return tree_zeros_like(state)
        result = self.linear_ops[0].apply(state, self.grid, t, self.params)
        for op in self.linear_ops[1:]:
            result = tree_add(result, op.apply(state, self.grid, t, self.params))
        return result

    def nonlinear_rhs(self, state: StateDict, t: float) -> StateDict:
        if len(self.nonlinear_ops) == 0:
            return tree_zeros_like(state)
        result = self.nonlinear_ops[0].apply(state, self.grid, t, self.params)
        for op in self.nonlinear_ops[1:]:
            result = tree_add(result, op.apply(state, self.grid, t, self.params))
        return result

    def rhs(self, state: StateDict, t: float) -> StateDict:
        """model.py:100-118: apply BCs, then linear + nonlinear parts."""
        state = self.apply_bcs(state, t)
        return tree_add(self.linear_rhs(state, t), self.nonlinear_rhs(state, t))

    def create_padded_field(self, init_fn: Optional[Callable] = None, fill_value: float = 0.0) -> torch.Tensor:
        dev = default_device()
        if isinstance(self.grid, Grid1D):
            if init_fn is not None:
                return torch.as_tensor(init_fn(self.grid.x_coords(include_ghost=True))).to(dev, self.dtype)
            return torch.full((self.grid.nx_total,), fill_value, dtype=self.dtype, device=dev)
        if init_fn is not None:
            X, Y = self.grid.meshgrid(include_ghost=True)
            return torch.as_tensor(init_fn(X, Y)).to(dev, self.dtype)
        return torch.full((self.grid.ny_total, self.grid.nx_total), fill_value, dtype=self.dtype, device=dev)

    def create_initial_state(self, init_fns: Optional[Dict[str, Callable]] = None,
                             fill_values: Optional[Dict[str, float]] = None) -> StateDict:
        init_fns = init_fns or {}
        fill_values = fill_values or {}
        state = {}
        for name in self.field_names:
            if name in init_fns:
                state[name] = self.create_padded_field(init_fn=init_fns[name])
            else:
                state[name] = self.create_padded_field(fill_value=fill_values.get(name, 0.0))
        return state

    def cfl_dt(self, state: Optional[StateDict] = None) -> float:
        dt_min = float("inf")
        for op in self.linear_ops:
            if op.dt_bound is not None:
                dt_min = min(dt_min, float(op.dt_bound(self.grid, self.params)))
        if state is not None:
            for op in self.nonlinear_ops:
                if op.dt_bound is not None:
                    dt_min = min(dt_min, float(op.dt_bound(self.grid, state, self.params)))
        return dt_min


def create_gray_scott_model(grid: Grid2D, Du: float = 0.16, Dv: float = 0.08, F: float = 0.04, k: float = 0.06,
                            bc_type: BCType = BCType.PERIODIC, dtype=torch.float64) -> MOLModel:
    """model.py:213-292: du/dt = Du lap u - u v^2 + F(1-u),
    dv/dt = Dv lap v + u v^2 - (F+k) v."""
    bc_spec = {"u": FieldBCSpec(kind=bc_type), "v": FieldBCSpec(kind=bc_type)}
    params = {"Du": Du, "Dv": Dv, "F": F, "k": k, "dtype": dtype}

    def diffusion_apply(state, grid, t, params):
        return {"u": params["Du"] * laplacian_2d(state["u"], grid),
                "v": params["Dv"] * laplacian_2d(state["v"], grid)}

    def diffusion_dt_bound(grid, params):
        return 0.25 * grid.min_dx2 / (max(params["Du"], params["Dv"]) + 1e-14)

    diffusion_op = LinearOp(name="gray_scott_diffusion", apply=diffusion_apply, dt_bound=diffusion_dt_bound)
    metadata = {"field_names": ["u", "v"], "pde_type": "reaction_diffusion", "stiff": True,
                "mjx_kernel": "gray_scott"}
    return MOLModel(grid=grid, bc_spec=bc_spec, params=params, linear_ops=(diffusion_op,),
                    nonlinear_ops=(gray_scott_reaction_op(),), metadata=metadata)


def create_gray_scott_periodic_fft(grid: Grid2D, Du: float = 0.16, Dv: float = 0.08, F: float = 0.04,
                                   k: float = 0.06, dtype=torch.float64):
    """model.py:529-584: (model, fft_cache, diffusivities)."""
    from .fft_solvers import create_fft_cache
    model = create_gray_scott_model(grid, Du=Du, Dv=Dv, F=F, k=k, bc_type=BCType.PERIODIC, dtype=dtype)
    fft_cache = create_fft_cache(grid, dtype)
    diffusivities = {"u": Du, "v": Dv}
    metadata = dict(model.metadata)
    metadata.update(fft_cache=fft_cache, diffusivities=diffusivities, supports_fft=True)
    model_fft = MOLModel(grid=model.grid, bc_spec=model.bc_spec, params=model.params,
                         linear_ops=model.linear_ops, nonlinear_ops=model.nonlinear_ops, metadata=metadata)
    return model_fft, fft_cache, diffusivities


def get_fft_cache(model: MOLModel):
    """model.py:645-647."""
    return model.metadata.get("fft_cache", None)


def get_diffusivities(model: MOLModel) -> Dict[str, float]:
    """model.py:650-667."""
    if "diffusivities" in model.metadata:
        return model.metadata["diffusivities"]
    out = {}
    for name in model.field_names:
        if f"D{name}" in model.params:
            out[name] = model.params[f"D{name}"]
        elif "D" in model.params:
            out[name] = model.params["D"]
        else:
            out[name] = 0.0
    return out
