"""
The advected tracer stack's layout and its open-edge inflow values, for callers that keep the stack on the device.

``stack_layout`` gives the positions of ``[T, (T²), S, (S²), tr…]`` from the model's switches, as
``core/model/driver.py:101-121`` (``_stack_layout``) does; ``stack_fields`` orders the fields that way
(``driver.py:124-136``); ``inflow_stack`` is ``OpenBoundary.inflow_stack`` (``core/components/open_boundary.py:353-380``):
the exterior values the advection reads on open edges, the stack itself where no edge is open or no inflow is given.
Both take NumPy arrays or torch tensors (then the result stays on their device); squares are formed as x·x, the
product NumPy's ``x**2`` forms.
"""
from __future__ import annotations

import numpy as np

try:
    import torch
except ImportError:                                                     # pragma: no cover (torch is a dependency)
    torch = None


def stack_layout(model) -> dict[str, int]:
    """Positions in the stack ``[T, (T²), S, (S²), tr…]`` (driver.py:101-121): 'T'; 'T2' unless the model is ALE
    without ``mix_diagnostic`` or ``variance_budget``; 'S' with ``salinity`` and 'S2' when it also has
    ``variance_budget`` or a ``tef_face``; 'tr' (the first passive tracer) for an ALE model with ``n_tracers``; 'n',
    the stack length.  ``model`` is any object with those attributes (missing ones read as off)."""
    lay, i = {"T": 0}, 1
    vb = bool(getattr(model, "variance_budget", False))
    tef = getattr(model, "tef_face", None) is not None
    ale = bool(getattr(model, "ale", False))
    if not ale or getattr(model, "mix_diagnostic", False) or vb:
        lay["T2"], i = i, i + 1
    if getattr(model, "salinity", False):
        lay["S"], i = i, i + 1
        if vb or tef:
            lay["S2"], i = i, i + 1
    ntr = int(getattr(model, "n_tracers", 0) or 0) if ale else 0
    if ntr:
        lay["tr"], i = i, i + ntr
    lay["n"] = i
    return lay


def _square(x):
    return x * x


def stack_fields(layout: dict[str, int], T, S, tr) -> list:
    """The fields in ``layout`` order: T, T², S, S², then the passive tracers ``tr`` (driver.py:124-136)."""
    out: list = [None] * layout["n"]
    out[layout["T"]] = T
    if "T2" in layout:
        out[layout["T2"]] = _square(T)
    if "S" in layout:
        out[layout["S"]] = S
    if "S2" in layout:
        out[layout["S2"]] = _square(S)
    if "tr" in layout:
        out[layout["tr"]:] = list(tr)
    return out


def _stack(parts):
    if torch is not None and isinstance(parts[0], torch.Tensor):
        return torch.stack(list(parts)).contiguous()
    return np.ascontiguousarray(np.stack(parts))


def inflow_stack(ob, c, T_in, tr_in, S_in=None, layout: dict[str, int] | None = None):
    """
    Inflow values of the advected stack ``c`` on open edges (open_boundary.py:353-380): T_in (and T_in² when the
    stack carries it), S_in (and S_in²) at their ``layout`` slots, then the passive tracers' inflow values ``tr_in``;
    ``c`` itself when ``T_in`` is None or ``ob.open`` is empty.  Without ``layout`` the stack is ``[T, (T²), tr…]``:
    its leading members are those of c that are not passive tracers (``ob.n_tracers`` of them when ``ob.ale``).
    ``ob``: an object with ``open`` (the open edge names) and, without ``layout``, ``ale`` and ``n_tracers``.
    """
    if T_in is None or not getattr(ob, "open", ()):
        return c
    if layout is not None:
        return _stack(stack_fields(layout, T_in, S_in, tr_in))
    ale = bool(getattr(ob, "ale", False))
    ntr = int(getattr(ob, "n_tracers", 0) or 0)
    lead = c.shape[0] - (ntr if ale else 0)
    parts = [T_in, _square(T_in)][:lead]
    if ale and ntr:
        parts += list(tr_in)
    return _stack(parts)
