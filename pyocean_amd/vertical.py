"""
Vertical regrid and conservative remap of the ALE coordinate on the device (HIP kernels of ``libocx.so``).

Mirrors ``core/vertical/remap.py`` (``remap_columns``, schemes PCM / PLM / PPM_H4), the z* and sigma rules of
``core/vertical/regrid.py`` and ``ALECoordinate.regrid`` / ``regrid_remap``
(core/components/vertical_coordinate.py:69-109), with the same argument meaning and errors; torch tensors in HBM
replace NumPy arrays.  Bit-identical to the reference's FP64 arithmetic.  The hybrid (HYCOM1) rule takes the layer
densities ``rho`` from the caller (the reference computes them from T and S with its equation of state).

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
_WS_HYB = Workspace()


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


def regrid(H, eta, active, h, rule: str, params: dict, check_columns: bool = True, rho=None):
    """
    New layer thicknesses of the rule at surface eta from layers h (n, ny, nx): (h_new, wet) with
    wet = active & H + eta > 0 & Σ h > 0; h kept where dry.  rule 'zstar' (params['dz_top'], nominal thicknesses
    surface first), 'sigma' (params['dsigma'], fractions bed first) or 'hybrid' (params['dz_top'], ['rho_target']
    (n+1, surface first), optional ['max_h_top'], ['d_sigma'], ['d_hybrid'] with ['dsigma'], ['keep_isopycnal']
    (default True); ``rho`` (n, ny, nx) the layer densities); params['h_min'] (default 1e-3 m).
    ``check_columns`` raises ValueError, as the reference does, for a wet column the rule cannot build (a host
    synchronisation).  core/components/vertical_coordinate.py:94-109, core/vertical/regrid.py:41-405.
    """
    if rule not in ("zstar", "sigma", "hybrid"):
        raise ValueError(f"unknown rule {rule!r}; use 'zstar', 'sigma' or 'hybrid'")
    required = {"zstar": ("dz_top",), "sigma": ("dsigma",), "hybrid": ("dz_top", "rho_target")}[rule]
    if rule == "hybrid" and float(params.get("d_hybrid", 0.0)) > 0.0:
        required += ("dsigma",)
    for key in required:
        if key not in params:
            raise ValueError(f"regrid_columns({rule!r}) needs params[{key!r}]")
    if rule == "hybrid" and rho is None:
        raise ValueError("regrid_columns('hybrid') needs the layer densities rho")
    n, ny, nx = h.shape
    hc = _f64(h, (n, ny, nx), "h")
    Hc, ec = _f64(H, (ny, nx), "H"), _f64(eta, (ny, nx), "eta")
    act = _mask(active, (ny, nx), "active")
    dev = hc.device

    def vec(key, default, size):
        v = torch.as_tensor(params.get(key, default), dtype=_F64, device=dev).contiguous()
        if v.shape != (size,):
            raise ValueError(f"params[{key!r}]: {size} values expected, got {tuple(v.shape)}")
        return v
    h_new = torch.empty_like(hc)
    wet = torch.empty((ny, nx), dtype=torch.uint8, device=dev)
    L = lib()
    bad = ctypes.c_longlong(-1)
    badp = ctypes.byref(bad) if check_columns else None
    h_min = float(params.get("h_min", H_MIN))
    if rule == "hybrid":
        rc_ = _f64(rho, (n, ny, nx), "rho")
        dz_top, dsigma = vec("dz_top", None, n), vec("dsigma", [1.0 / n] * n, n)
        targets, max_h = vec("rho_target", None, n + 1), vec("max_h_top", [float("inf")] * n, n)
        ws = _WS_HYB.get(int(L.ocx_regrid_hybrid_workspace_bytes(n, ny * nx)))
        rc = L.ocx_regrid_hybrid_f64(_ptr(Hc), _ptr(ec), _ptr(act), _ptr(hc), _ptr(rc_), n, ny * nx, _ptr(dz_top),
                                     _ptr(dsigma), _ptr(targets), _ptr(max_h), float(params.get("d_sigma", 0.0)),
                                     float(params.get("d_hybrid", 0.0)), h_min,
                                     int(bool(params.get("keep_isopycnal", True))), _ptr(h_new), _ptr(wet), _ptr(ws),
                                     ws.numel(), badp, _stream())
    else:
        dz = vec("dz_top" if rule == "zstar" else "dsigma", None, n)
        ws = _WS.get(int(L.ocx_regrid_workspace_bytes()))
        rc = L.ocx_regrid_f64(_ptr(Hc), _ptr(ec), _ptr(act), _ptr(hc), n, ny * nx, 0 if rule == "zstar" else 1,
                              _ptr(dz), h_min, _ptr(h_new), _ptr(wet), _ptr(ws), ws.numel(), badp, _stream())
    if rc == -4:
        b = bad.value
        raise ValueError(f"{L.ocx_last_error().decode()} (column j={b // nx}, i={b % nx})")
    check(rc, "regrid")
    return h_new, wet


class VerticalCoordinate:
    """
    The ALE vertical coordinate's regrid + remap step (``ALECoordinate``, core/components/vertical_coordinate.py:
    27-109) for the z*, sigma and hybrid rules: H, active (ny, nx); rule; params as the reference's (see
    ``regrid``); remap_code PCM / PLM / PPM_H4; ``eos`` (``pyocean_amd.eos``) and ``p_ref`` [Pa] give the layer
    densities of the hybrid rule.
    """

    def __init__(self, *, H, active, rule: str, params: dict, remap_code: int = PPM_H4, eos=None,
                 p_ref: float = 2e7):
        if rule not in ("zstar", "sigma", "hybrid"):
            raise ValueError(f"unknown rule {rule!r}; use 'zstar', 'sigma' or 'hybrid'")
        self.H, self.active, self.rule, self.params, self.remap_code = H, active, rule, params, int(remap_code)
        self.eos, self.p_ref = eos, float(p_ref)

    def regrid_density(self, T, S=None):
        """Layer densities the hybrid rule follows [kg m^-3]: the linear EOS's density, else the potential density
        at p_ref (core/components/vertical_coordinate.py:80-100), unmodified."""
        if self.eos is None:
            raise ValueError("regrid_density needs an equation of state (eos=...)")
        if self.eos.linear:
            return self.eos.density(T, S)
        return self.eos.potential_density(T, S, self.p_ref)

    def _rho(self, T, S, rho):
        if self.rule != "hybrid" or rho is not None:
            return rho
        if T is None:
            raise ValueError("the hybrid rule needs T (and S) or the layer densities rho")
        return self.regrid_density(T, S)

    def regrid(self, eta, h, T=None, S=None, rho=None):
        """New layer thicknesses of the rule at surface eta from layers h, h kept where dry
        (core/components/vertical_coordinate.py:102-117); the hybrid rule follows regrid_density(T, S) or
        the given densities ``rho``."""
        return regrid(self.H, eta, self.active, h, self.rule, self.params, rho=self._rho(T, S, rho))[0]

    def regrid_remap(self, eta, h_src, fields, T=None, S=None, rho=None):
        """(h_new, remapped): ``fields`` a list of (n, ny, nx) tensors (a list comes back) or one stack
        (nf, n, ny, nx) (a stack comes back, one launch).  The hybrid rule follows regrid_density(T, S), or the
        given layer densities ``rho`` (core/components/vertical_coordinate.py:69-78)."""
        h_new, wet = regrid(self.H, eta, self.active, h_src, self.rule, self.params, rho=self._rho(T, S, rho))
        if torch.is_tensor(fields):
            return h_new, remap_columns(h_src, fields, h_new, self.remap_code, wet)
        return h_new, [remap_columns(h_src, f, h_new, self.remap_code, wet) for f in fields]
