"""
The real-freshwater substep of the tracers on the device (``libocx.so``): the substep that follows the advection and
precedes the vertical diffusion.

Mirrors ``freshwater_volumes`` and ``apply_freshwater`` (core/components/freshwater.py:31-122) with the same
argument meaning and results, torch float64 GPU tensors in place of NumPy arrays; bit-identical to the reference's
FP64 arithmetic.  Freshwater carries S = 0 and the top layer's temperature: salt-like members (S, S^2, passive
tracers) keep their content, T and T^2 their value.
"""
from __future__ import annotations

import ctypes

import torch

from ._native import check, lib
from .ops import Workspace, _ptr, _stream

_F64 = torch.float64
_WS = Workspace()


def _f64(t, shape, name):
    if not torch.is_tensor(t) or t.dtype != _F64 or not t.is_cuda:
        raise ValueError(f"{name}: a float64 GPU tensor")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)}")
    return t.contiguous()


def freshwater_volumes(V, dV, floor, *, scale: float = 1.0, area=None):
    """
    (V_new, n_limited): the freshwater volume dV (ny, nx) added to the top layer (k = N-1) of V (N, ny, nx); a
    removal that would take the top layer below ``floor`` (a number or (ny, nx)) is taken from the layers below,
    top-down, each first kept at the floor, then down to zero.  n_limited counts those columns (a host
    synchronisation).  core/components/freshwater.py:31-58.  ``scale`` multiplies dV first (the layer transport
    passes dt and fw); with ``area`` the result is (V_new, V_new A, n_limited).
    """
    n, ny, nx = V.shape
    Vc, d = _f64(V, (n, ny, nx), "V"), _f64(dV, (ny, nx), "dV")
    fc = _f64(floor, (ny, nx), "floor") if torch.is_tensor(floor) else None
    A = _f64(area, (ny, nx), "area") if area is not None else None
    out = torch.empty_like(Vc)
    vout = torch.empty_like(Vc) if A is not None else None
    L = lib()
    ws = _WS.get(int(L.ocx_freshwater_workspace_bytes()))
    cnt = ctypes.c_longlong(0)
    check(L.ocx_freshwater_volumes_f64(_ptr(Vc), _ptr(d), float(scale), _ptr(fc),
                                       0.0 if fc is not None else float(floor), _ptr(A), n, ny, nx, _ptr(out),
                                       _ptr(vout), _ptr(ws), ws.numel(), ctypes.byref(cnt), _stream()),
          "freshwater_volumes")
    if A is not None:
        return out, vout, int(cnt.value)
    return out, int(cnt.value)


def apply_freshwater(stack, layout: dict, V_adv, V_fw, fw_area_dt, *, dt: float):
    """
    The freshwater substep on the tracer stack (nt, N, ny, nx), in place (contiguous), returning
    (stack, n_limited, var_surf): ``layout`` the stack positions ("T", optional "T2", optional "S", "n"); V_adv the
    layer volumes after the advection, V_fw those after the freshwater (N, ny, nx); fw_area_dt = dt A fw (ny, nx).
    var_surf (N, ny, nx) the salinity-variance source, None without "S".  core/components/freshwater.py:61-122.
    """
    nt, n, ny, nx = stack.shape
    if stack.dtype != _F64 or not stack.is_cuda or not stack.is_contiguous():
        raise ValueError("stack: a contiguous float64 GPU tensor")
    if int(layout["n"]) != nt:
        raise ValueError(f"layout['n'] = {layout['n']} but the stack holds {nt} members")
    Va, Vf = _f64(V_adv, (n, ny, nx), "V_adv"), _f64(V_fw, (n, ny, nx), "V_fw")
    fwa = _f64(fw_area_dt, (ny, nx), "fw_area_dt")
    t2 = layout.get("T2")
    s_index = layout.get("S")
    var = torch.empty((n, ny, nx), dtype=_F64, device=stack.device) if s_index is not None else None
    L = lib()
    ws = _WS.get(int(L.ocx_freshwater_workspace_bytes()))
    cnt = ctypes.c_longlong(0)
    check(L.ocx_apply_freshwater_f64(_ptr(stack), nt, n, ny, nx, int(layout["T"]), -1 if t2 is None else int(t2),
                                     -1 if s_index is None else int(s_index), _ptr(Va), _ptr(Vf), _ptr(fwa), float(dt),
                                     _ptr(var), _ptr(ws), ws.numel(), ctypes.byref(cnt), _stream()),
          "apply_freshwater")
    return stack, int(cnt.value), var
