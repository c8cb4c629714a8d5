"""
Vertical regrid and conservative remap of the ALE coordinate on the device (HIP kernels of ``libocx.so``).

Mirrors ``core/vertical/remap.py`` (``remap_columns``, schemes PCM / PLM / PPM_H4), the z* and sigma rules of
``core/vertical/regrid.py`` and ``ALECoordinate.regrid`` / ``regrid_remap``
(core/components/vertical_coordinate.py:69-109), with the same argument meaning and errors; torch tensors in HBM
replace NumPy arrays.  Bit-identical to the reference's FP64 arithmetic.  The hybrid (HYCOM1) rule is not offered
and raises ``NotImplementedError``.

A field stack (nf, n, ny, nx) - the tracer stack C, say - is remapped in one launch without copies: the thickness
reads are shared by the fields of a column.
"""
from __future__ import annotations

import ctypes

import torch

from ._native import check, lib
from .ops import Workspace, _ptr, _stream

PCM, PLM, PPM_H4 = 0, 1, 2
SCHEMES = {"pcm": PCM, "plm": PLM, "ppm_h4": PPM_H4}
H_MIN = 1e-3            # minimum layer thickness after regridding [m] (core/vertical/regrid.py:40)
_F64 = torch.float64
_WS = Workspace()


def _f64(t, shape, name):
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)}")
    if t.dtype != _F64 or not t.is_cuda:
        raise ValueError(f"{name}: a float64 GPU tensor")
    return t.contiguous()


def _mask(t, shape, name):
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)}")
    if t.dtype == torch.bool:
        t = t.contiguous().view(torch.uint8)
    if t.dtype != torch.uint8 or not t.is_cuda:
        raise ValueError(f"{name}: a bool or uint8 GPU tensor")
    return t.contiguous()


def remap_columns(h_src, f_src, h_tgt, scheme: int, wet, out=None):
    """
    Remap every wet column of f_src (N0, ny, nx) - or a stack of fields (nf, N0, ny, nx) - from thicknesses h_src
    (N0, ny, nx) to h_tgt (N1, ny, nx) [m].  Dry columns: f_src copied when N0 == N1, else NaN.
    core/vertical/remap.py:412-427 (remap_columns), :259-383 (remap_column).
    """
    n0, *cols = h_src.shape
    n1 = h_tgt.shape[0]
    stack = f_src.dim() == h_src.dim() + 1
    nf = f_src.shape[0] if stack else 1
    hs = _f64(h_src, (n0, *cols), "h_src")
    ht = _f64(h_tgt, (n1, *cols), "h_tgt")
    f = _f64(f_src, ((nf,) if stack else ()) + (n0, *cols), "f_src")
    w = _mask(wet, tuple(cols), "wet")
    M = hs.numel() // n0
    shape = ((nf,) if stack else ()) + (n1, *cols)
    o = torch.empty(shape, dtype=_F64, device=hs.device) if out is None else _f64(out, shape, "out")
    if int(scheme) not in (PCM, PLM, PPM_H4):
        raise ValueError(f"scheme {scheme}: PCM, PLM or PPM_H4")
    check(lib().ocx_remap_columns_f64(_ptr(hs), _ptr(f), _ptr(ht), _ptr(w), n0, n1, M, nf, int(scheme), _ptr(o),
                                      _stream()), "remap_columns")
    return o


def regrid(H, eta, active, h, rule: str, params: dict, check_columns: bool = True):
    """
    New layer thicknesses of the rule at surface eta from layers h (n, ny, nx): (h_new, wet) with
    wet = active & H + eta > 0 & Σ h > 0; h kept where dry.  rule 'zstar' (params['dz_top'], nominal thicknesses
    surface first) or 'sigma' (params['dsigma'], fractions bed first); params['h_min'] (default 1e-3 m).
    ``check_columns`` raises ValueError, as the reference does, for a wet column the rule cannot build (a host
    synchronisation).  core/components/vertical_coordinate.py:94-109, core/vertical/regrid.py:41-95, 360-405.
    """
    if rule == "hybrid":
        raise NotImplementedError("the hybrid (HYCOM1) rule is not offered on the device")
    if rule not in ("zstar", "sigma"):
        raise ValueError(f"unknown rule {rule!r}; use 'zstar', 'sigma' or 'hybrid'")
    key = "dz_top" if rule == "zstar" else "dsigma"
    if key not in params:
        raise ValueError(f"regrid_columns({rule!r}) needs params[{key!r}]")
    n, ny, nx = h.shape
    hc = _f64(h, (n, ny, nx), "h")
    Hc, ec = _f64(H, (ny, nx), "H"), _f64(eta, (ny, nx), "eta")
    act = _mask(active, (ny, nx), "active")
    dz = torch.as_tensor(params[key], dtype=_F64, device=hc.device).contiguous()
    if dz.shape != (n,):
        raise ValueError(f"params[{key!r}]: {n} values expected, got {tuple(dz.shape)}")
    h_new = torch.empty_like(hc)
    wet = torch.empty((ny, nx), dtype=torch.uint8, device=hc.device)
    L = lib()
    ws = _WS.get(int(L.ocx_regrid_workspace_bytes()))
    bad = ctypes.c_longlong(-1)
    rc = L.ocx_regrid_f64(_ptr(Hc), _ptr(ec), _ptr(act), _ptr(hc), n, ny * nx, 0 if rule == "zstar" else 1, _ptr(dz),
                          float(params.get("h_min", H_MIN)), _ptr(h_new), _ptr(wet), _ptr(ws), ws.numel(),
                          ctypes.byref(bad) if check_columns else None, _stream())
    if rc == -4:
        b = bad.value
        raise ValueError(f"wet columns need finite H and D = H + eta > 0{' and H > 0 for z*' if rule == 'zstar' else ''}"
                         f" (column j={b // nx}, i={b % nx})")
    check(rc, "regrid")
    return h_new, wet


class VerticalCoordinate:
    """
    The ALE vertical coordinate's regrid + remap step (``ALECoordinate.regrid_remap``,
    core/components/vertical_coordinate.py:69-79) for the z* and sigma rules: H, active (ny, nx); rule; params as
    the reference's (dz_top / dsigma, h_min); remap_code PCM / PLM / PPM_H4.
    """

    def __init__(self, *, H, active, rule: str, params: dict, remap_code: int = PPM_H4):
        if rule == "hybrid":
            raise NotImplementedError("the hybrid (HYCOM1) rule is not offered on the device")
        self.H, self.active, self.rule, self.params, self.remap_code = H, active, rule, params, int(remap_code)

    def regrid_remap(self, eta, h_src, fields, T=None, S=None):
        """(h_new, remapped): ``fields`` a list of (n, ny, nx) tensors (a list comes back) or one stack
        (nf, n, ny, nx) (a stack comes back, one launch).  T and S are read by the hybrid rule only."""
        h_new, wet = regrid(self.H, eta, self.active, h_src, self.rule, self.params)
        if torch.is_tensor(fields):
            return h_new, remap_columns(h_src, fields, h_new, self.remap_code, wet)
        return h_new, [remap_columns(h_src, f, h_new, self.remap_code, wet) for f in fields]
