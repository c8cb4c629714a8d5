"""
The pieces of pyOCEAN's component contract the advection component exposes (``core/model/contract.py:27-55``):
its role and the configuration-issue record returned by ``check``.
"""
from __future__ import annotations

from dataclasses import dataclass
from enum import Enum
from typing import Literal


class Role(str, Enum):
    """Roles of the components, one per place in the fixed step order (contract.py:27-43)."""
    EOS = "eos"
    VERTICAL_COORDINATE = "vertical_coordinate"
    PGF = "pgf"
    MOMENTUM = "momentum"
    TIME_STEPPER = "time_stepper"
    FREE_SURFACE = "free_surface"
    CLOSURE = "closure"
    BOTTOM_DRAG = "bottom_drag"
    LAYER_TRANSPORT = "layer_transport"
    TRACER_ADVECTION = "tracer_advection"
    TRACER_DIFFUSION = "tracer_diffusion"
    OPEN_BOUNDARY = "open_boundary"
    INTERIOR_SPONGE = "interior_sponge"
    SURFACE_FLUXES = "surface_fluxes"
    DIAGNOSTICS = "diagnostics"


@dataclass(frozen=True)
class Issue:
    """One configuration problem found before the first step (contract.py:46-51)."""
    level: Literal["error", "warning"]
    component: str
    message: str
