"""Boundary-condition specification (moljax/core/bc.py:26-78).

Ghost cells are filled by the kernels that produce the fields (the fused
spectral and stepping kernels write the periodic ring directly).  The
standalone ``apply_bc`` of the reference is provided for padded states that
come from user code.
"""
from dataclasses import dataclass
from enum import IntEnum
from typing import Callable, Dict, Optional


class BCType(IntEnum):
    """moljax/core/bc.py:26-30."""
    PERIODIC = 0
    DIRICHLET = 1
    NEUMANN = 2


@dataclass(frozen=True)
class FieldBCSpec:
    """moljax/core/bc.py:33-78."""
    kind: BCType
    value: Optional[Callable] = None
    flux: Optional[Callable] = None

    @classmethod
    def periodic(cls) -> "FieldBCSpec":
        return cls(kind=BCType.PERIODIC)

    @classmethod
    def dirichlet(cls, value: Optional[Callable] = None) -> "FieldBCSpec":
        return cls(kind=BCType.DIRICHLET, value=value)

    @classmethod
    def neumann(cls, flux: Optional[Callable] = None) -> "FieldBCSpec":
        return cls(kind=BCType.NEUMANN, flux=flux)


BCSpec = Dict[str, FieldBCSpec]


def create_default_bc_spec(field_names: list, bc_type: BCType = BCType.PERIODIC) -> BCSpec:
    """moljax/core/bc.py:407-421."""
    return {name: FieldBCSpec(kind=bc_type) for name in field_names}
