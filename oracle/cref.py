"""
Build recipe and ctypes wrapper of the C loop oracle ``loops.c`` (test infrastructure only; see ``oracle/__init__``).

``build()`` compiles ``oracle/_build/liboracle.so`` with gcc (``-O3 -ffp-contract=off -fopenmp``).  The library
is git-ignored but travels with the tree, so the CPU baseline can be timed where no compiler run is wanted.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
SRC = HERE / "loops.c"
SRCS = [SRC, HERE / "remap.c"]


def build(verbose: bool = False) -> Path:
    """Compile the C oracle (always; a few seconds)."""
    LIB.parent.mkdir(exist_ok=True)
    tmp = LIB.with_suffix(f".so.tmp{os.getpid()}")
    cmd = ["gcc", "-O3", "-march=x86-64-v2", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC",
           "-shared", "-std=c11", *map(str, SRCS), "-o", str(tmp), "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        P, I, D, LL = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_longlong
        L.oc_set_threads.argtypes = [I]
        L.oc_advect_fct.argtypes = [P, P, P, P, P, P, P, I, I, I, I, D, I, I, I, I, P, P, P, P]
        L.oc_advect_fct.restype = I
        L.oc_advect_rk3.argtypes = [P, P, P, P, P, P, P, P, I, I, I, I, D, I, I, P]
        L.oc_advect_rk3.restype = I
        L.oc_faces_h.argtypes = [P, P, P, P, P, I, I, I, I, I, I, P, P]
        L.oc_fct_coeff.argtypes = [P, P, P, P, P, P, D, I, I, I, I, P, P]
        L.oc_advect_div.argtypes = [P, P, P, P, P, P, I, I, I, I, I, I, I, P]
        L.oc_reconstruct.argtypes = [P, P, I, I, P]
        L.oc_remap_column.argtypes = [P, P, I, P, I, I, P]
        L.oc_remap_columns.argtypes = [P, P, I, P, I, LL, I, I, P, P]
        L.oc_zstar_column.argtypes = [D, D, P, I, D, P]
        L.oc_sigma_column.argtypes = [D, D, P, I, D, P]
        L.oc_regrid.argtypes = [P, P, P, P, I, LL, I, P, D, P, P]
        L.oc_hycom1_column.argtypes = [D, D, P, P, I, P, P, P, D, I, P]
        L.oc_window_shallow_sigma.argtypes = [P, P, P, D, D, D, D, D, I, P]
        L.oc_regrid_hybrid.argtypes = [P, P, P, P, P, I, LL, P, P, P, P, D, D, D, I, P, P]
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    """OpenMP team size of the C loops."""
    lib().oc_set_threads(int(n))


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _u8(a, shape):
    return np.ascontiguousarray(np.broadcast_to(a, shape), dtype=np.uint8)


def advect_fct(C, V0, V1, Qx, Qy, ok, dt, inflow=None, open_mask=0, z_variant=True, unlimited=False,
               flux_face_i=None, want_faces=False):
    """One FCT step by the C loops (same meaning as ``numpy_ref.advect_fct``).  Returns out, or a dict with
    ``out``, ``fx_i`` and the effective face fluxes ``Gx``/``Gy`` when asked for."""
    C = _f64(C)
    nt, n, ny, nx = C.shape
    args = [_f64(V0), _f64(V1), _f64(Qx), _f64(Qy), _u8(ok, (n, ny, nx))]
    ext = None if inflow is None else _f64(inflow)
    out = np.empty_like(C)
    fxi = np.empty((nt, n, ny)) if flux_face_i is not None else None
    Gx = np.empty((nt, n, ny, nx + 1)) if want_faces else None
    Gy = np.empty((nt, n, ny + 1, nx)) if want_faces else None
    rc = lib().oc_advect_fct(_p(C), *map(_p, args), _p(ext), nt, n, ny, nx, float(dt), int(open_mask),
                             int(bool(z_variant)), int(bool(unlimited)), -1 if flux_face_i is None else int(flux_face_i),
                             _p(out), _p(fxi), _p(Gx), _p(Gy))
    if rc != 0:
        raise MemoryError("oracle: allocation failed")
    if flux_face_i is None and not want_faces:
        return out
    return {"out": out, "fx_i": fxi, "Gx": Gx, "Gy": Gy}


def advect_rk3(C, V0, V1, Qx, Qy, Fz, ok, dt, inflow=None, open_mask=0, z_variant=True):
    """One SSP-RK3 step by the C loops (same meaning as ``numpy_ref.advect_rk3``)."""
    C = _f64(C)
    nt, n, ny, nx = C.shape
    args = [_f64(V0), _f64(V1), _f64(Qx), _f64(Qy), _f64(Fz), _u8(ok, (n, ny, nx))]
    ext = None if inflow is None else _f64(inflow)
    out = np.empty_like(C)
    rc = lib().oc_advect_rk3(_p(C), *map(_p, args), _p(ext), nt, n, ny, nx, float(dt), int(open_mask),
                             int(bool(z_variant)), _p(out))
    if rc != 0:
        raise MemoryError("oracle: allocation failed")
    return out


def rhs(c, Qx, Qy, Fz, ok, ext, open_mask=0, z_variant=True, vert=None):
    """The advection RHS by the C loops (same meaning as ``numpy_ref.rhs``; ``vert`` None: any Fz != 0)."""
    c = _f64(c)
    nt, n, ny, nx = c.shape
    Fz = _f64(Fz)
    if vert is None:
        vert = bool(np.any(Fz != 0.0))
    out = np.empty_like(c)
    lib().oc_advect_div(_p(c), _p(_f64(Qx)), _p(_f64(Qy)), _p(Fz), _p(_u8(ok, (n, ny, nx))), _p(_f64(ext)), nt, n,
                        ny, nx, int(open_mask), int(bool(z_variant)), int(bool(vert)), _p(out))
    return out


def faces_h(c, Qx, Qy, ok, ext, open_mask=0, z_variant=True):
    c = _f64(c)
    nt, n, ny, nx = c.shape
    fx = np.zeros((nt, n, ny, nx + 1))
    fy = np.zeros((nt, n, ny + 1, nx))
    lib().oc_faces_h(_p(c), _p(_f64(Qx)), _p(_f64(Qy)), _p(_u8(ok, (n, ny, nx))), _p(_f64(ext)), nt, n, ny, nx,
                     int(open_mask), int(bool(z_variant)), _p(fx), _p(fy))
    return fx, fy


def fct_coefficients(c0, cL, V1, Ax, Ay, ok, dt):
    c0 = _f64(c0)
    nt, n, ny, nx = c0.shape
    Cx = np.zeros((nt, n, ny, nx + 1))
    Cy = np.zeros((nt, n, ny + 1, nx))
    lib().oc_fct_coeff(_p(c0), _p(_f64(cL)), _p(_f64(V1)), _p(_f64(Ax)), _p(_f64(Ay)), _p(_u8(ok, (n, ny, nx))),
                       float(dt), nt, n, ny, nx, _p(Cx), _p(Cy))
    return Cx, Cy


# ------------------------------------------------------------------------------------------------ vertical remap
PCM, PLM, PPM_H4 = 0, 1, 2
SCHEMES = {"pcm": PCM, "plm": PLM, "ppm_h4": PPM_H4}


def reconstruct(h, u, scheme):
    """(n, 3) polynomial coefficients of every cell (``reconstruct``, core/vertical/remap.py:174-245)."""
    h, u = _f64(h), _f64(u)
    coef = np.empty((h.size, 3))
    lib().oc_reconstruct(_p(h), _p(u), h.size, int(scheme), _p(coef))
    return coef


def remap_column(h_src, u_src, h_tgt, scheme):
    """One column (``remap_column``, core/vertical/remap.py:259-383)."""
    h_src, u_src, h_tgt = _f64(h_src), _f64(u_src), _f64(h_tgt)
    out = np.empty(h_tgt.size)
    lib().oc_remap_column(_p(h_src), _p(u_src), h_src.size, _p(h_tgt), h_tgt.size, int(scheme), _p(out))
    return out


def remap_columns(h_src, f_src, h_tgt, scheme, wet):
    """Every wet column (``remap_columns``, core/vertical/remap.py:386-427): h_src (n0, ...), f_src (n0, ...) or a
    stack (nf, n0, ...), h_tgt (n1, ...), wet (...).  Dry columns: copied when n0 == n1, else NaN."""
    h_src, f, h_tgt = _f64(h_src), _f64(f_src), _f64(h_tgt)
    n0, n1 = h_src.shape[0], h_tgt.shape[0]
    M = int(np.prod(h_src.shape[1:]))
    stack = f.ndim == h_src.ndim + 1
    nf = f.shape[0] if stack else 1
    out = np.empty((nf, n1) + h_src.shape[1:])
    lib().oc_remap_columns(_p(h_src), _p(f), n0, _p(h_tgt), n1, M, nf, int(scheme), _p(_u8(wet, h_src.shape[1:])),
                           _p(out))
    return out if stack else out[0]


def zstar_column(H, eta, dz_top, h_min=1e-3):
    dz = _f64(dz_top)
    zw = np.empty(dz.size + 1)
    lib().oc_zstar_column(float(H), float(eta), _p(dz), dz.size, float(h_min), _p(zw))
    return zw


def sigma_column(H, eta, dsigma, h_min=1e-3):
    ds = _f64(dsigma)
    zw = np.empty(ds.size + 1)
    lib().oc_sigma_column(float(H), float(eta), _p(ds), ds.size, float(h_min), _p(zw))
    return zw


def regrid(H, eta, active, h, rule, dz, h_min=1e-3):
    """``ALECoordinate.regrid`` for rule 'zstar' (dz = nominal thicknesses, surface first) or 'sigma' (dz = the
    fractions, bed first): (h_new, wet) with wet = active & H + eta > 0 & Σ h > 0 (vertical_coordinate.py:69-109)."""
    h = _f64(h)
    n = h.shape[0]
    M = int(np.prod(h.shape[1:]))
    h_new = np.empty_like(h)
    wet = np.empty(h.shape[1:], np.uint8)
    lib().oc_regrid(_p(_f64(H)), _p(_f64(eta)), _p(_u8(active, h.shape[1:])), _p(h), n, M,
                    {"zstar": 0, "sigma": 1}[rule], _p(_f64(dz)), float(h_min), _p(h_new), _p(wet))
    return h_new, wet.astype(bool)


def hycom1_column(H, eta, h_bed, rho_bed, rho_target, dz_top, max_h_top=None, h_min=1e-3, keep=False):
    """Hybrid interfaces of one column, bed first (``hycom1_column``, core/vertical/regrid.py:206-259)."""
    h, r = _f64(h_bed), _f64(rho_bed)
    n = h.size
    mh = _f64(np.full(n, np.inf) if max_h_top is None else max_h_top)
    zw = np.empty(n + 1)
    lib().oc_hycom1_column(float(H), float(eta), _p(h), _p(r), n, _p(_f64(rho_target)), _p(_f64(dz_top)), _p(mh),
                           float(h_min), int(bool(keep)), _p(zw))
    return zw


def window_shallow_sigma(zw_hyb, zw_sig, zw_zs, H, D, d_sigma, d_hybrid, h_min=1e-3):
    """``window_shallow_sigma`` (core/vertical/regrid.py:271-303)."""
    a, b, c = _f64(zw_hyb), _f64(zw_sig), _f64(zw_zs)
    out = np.empty_like(a)
    lib().oc_window_shallow_sigma(_p(a), _p(b), _p(c), float(H), float(D), float(d_sigma), float(d_hybrid),
                                  float(h_min), a.size - 1, _p(out))
    return out


def regrid_hybrid(H, eta, active, h, rho, params):
    """``ALECoordinate.regrid`` with the hybrid rule: (h_new, wet) (regrid.py:325-357, vertical_coordinate.py:94-109)."""
    h = _f64(h)
    n = h.shape[0]
    M = int(np.prod(h.shape[1:]))
    h_new = np.empty_like(h)
    wet = np.empty(h.shape[1:], np.uint8)
    f = lambda k, d: _f64(params.get(k, d))                                      # noqa: E731
    lib().oc_regrid_hybrid(_p(_f64(H)), _p(_f64(eta)), _p(_u8(active, h.shape[1:])), _p(h), _p(_f64(rho)), n, M,
                           _p(f("dz_top", np.ones(n))), _p(f("dsigma", np.full(n, 1.0 / n))),
                           _p(f("rho_target", np.zeros(n + 1))), _p(f("max_h_top", np.full(n, np.inf))),
                           float(params.get("d_sigma", 0.0)), float(params.get("d_hybrid", 0.0)),
                           float(params.get("h_min", 1e-3)), int(bool(params.get("keep_isopycnal", True))),
                           _p(h_new), _p(wet))
    return h_new, wet.astype(bool)
