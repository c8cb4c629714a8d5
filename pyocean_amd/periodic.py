"""
Periodic horizontal axes (``core/model/periodic.py:26-136``): which axes wrap, and the tracer halo width of the
pad-crop that the advection runs inside.  The wrapping itself runs on the GPU (``ops.periodic_pad``).
"""
from __future__ import annotations

AXES = ("x", "y")
TRACER_HALO = 13        # SSP-RK3 WENO5 (3 x 3 cells) + FCT donor/bounds/coefficients (3) + 1 spare


class Periodic:
    """Which horizontal axes wrap (``axes``: 'x', 'y' or both); falsy when none does (periodic.py:35-58)."""

    def __init__(self, axes=()) -> None:
        if isinstance(axes, str):
            axes = (axes,)
        axes = tuple(axes or ())
        bad = [a for a in axes if a not in AXES]
        if bad:
            raise ValueError(f"unknown periodic axes {bad}; use 'x', 'y' or both")
        self.x, self.y = "x" in axes, "y" in axes
        self.axes = tuple(a for a in AXES if a in axes)

    def __bool__(self) -> bool:
        return self.x or self.y

    def __repr__(self) -> str:
        return f"Periodic({self.axes})"


NONE = Periodic()
