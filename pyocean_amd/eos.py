"""
Equations of state for the layer densities of the hybrid regrid rule, on the device (``libocx.so``).

Mirrors ``LinearEOS`` and ``WrightFullEOS`` (core/components/eos.py:45-143) and ``make_eos`` (:161-178): the same
fields, defaults, names and errors, with torch float64 GPU tensors in place of NumPy arrays.  ``density`` runs one
elementwise kernel whose FP64 arithmetic equals the reference's bit for bit.  The TEOS-10 polynomial (Roquet et al.
2015) is not offered; the reference builds it only where the ``gsw`` package is installed.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from ._native import check, lib
from .ops import _ptr, _stream

EOS_LINEAR, EOS_WRIGHT_FULL = 0, 1


def _cells(x, like, name):
    if x is None:
        return None
    if not torch.is_tensor(x) or x.dtype != torch.float64 or not x.is_cuda:
        raise ValueError(f"{name}: a float64 GPU tensor")
    if tuple(x.shape) != tuple(like.shape):
        raise ValueError(f"{name}: shape {tuple(x.shape)}, expected {tuple(like.shape)}")
    return x.contiguous()


def _density(kind, T, S, p, linear=None):
    if not torch.is_tensor(T) or T.dtype != torch.float64 or not T.is_cuda:
        raise ValueError("T: a float64 GPU tensor")
    Tc = T.contiguous()
    Sc = _cells(S, Tc, "S")
    p_arr = _cells(p, Tc, "p") if torch.is_tensor(p) else None
    p_val = 0.0 if (p is None or p_arr is not None) else float(p)
    rho = torch.empty_like(Tc)
    if Tc.numel() == 0:
        return rho
    lin = (ctypes.c_double * 5)(*linear) if linear is not None else None
    check(lib().ocx_regrid_density_f64(kind, _ptr(Tc), _ptr(Sc), _ptr(p_arr), p_val,
                                       ctypes.cast(lin, ctypes.c_void_p) if lin is not None else None,
                                       Tc.numel(), _ptr(rho), _stream()), "regrid_density")
    return rho


@dataclass(frozen=True)
class LinearEOS:
    """rho = rho0 - t_coef (T - t_ref) + s_coef (S - s_ref) [kg m^-3] (Vallis 2006 eq. 1.156); p ignored."""

    rho0: float = 1000.0
    t_coef: float = 0.2
    t_ref: float = 0.0
    s_coef: float = 0.0
    s_ref: float = 35.0

    name = "linear"
    linear = True
    ts_kind = "theta_sp"

    def density(self, T, S=None, p=None):
        """core/components/eos.py:62-69: without S the temperature-only expression."""
        return _density(EOS_LINEAR, T, S, None, (self.rho0, self.t_coef, self.t_ref, self.s_coef, self.s_ref))

    def potential_density(self, T, S, p_ref=None):
        """The in situ density (the linear EOS does not depend on p)."""
        return self.density(T, S)


@dataclass(frozen=True)
class WrightFullEOS:
    """Wright (1997) full-range EOS on (theta, S_P) (MOM6 WRIGHT_FULL), p the sea pressure [Pa]."""

    rho0: float = 1035.0

    name = "wright_full"
    linear = False
    ts_kind = "theta_sp"

    def density(self, T, S=None, p=None):
        """rho(T, S, p) [kg m^-3]; p a tensor of T's shape or a number, 0 (the surface) by default
        (core/components/eos.py:104-108, physics/eos/wright_full.py:80-89)."""
        if S is None:
            raise ValueError(f"{self.name} EOS needs salinity")
        return _density(EOS_WRIGHT_FULL, T, S, 0.0 if p is None else p)

    def potential_density(self, T, S, p_ref):
        """Density referenced to p_ref [Pa] (sigma_2: p_ref = 2e7) (core/components/eos.py:118-120)."""
        return self.density(T, S, p_ref)


def make_eos(name: str, rho0: float | None = None, **linear_params):
    """EOS by name, as core/components/eos.py:161-178: 'linear' (t_coef, t_ref, s_coef, s_ref) or 'wright_full'."""
    kw = {} if rho0 is None else {"rho0": rho0}
    if name == "linear":
        return LinearEOS(**kw, **linear_params)
    if linear_params:
        raise ValueError(f"linear_params {sorted(linear_params)} only apply to the 'linear' EOS")
    if name == "wright_full":
        return WrightFullEOS(**kw)
    if name == "teos10":
        raise NotImplementedError("eos='teos10' is not offered on the device; pass the densities (rho) instead")
    raise ValueError(f"unknown EOS {name!r}: expected 'linear', 'wright_full' or 'teos10'")
