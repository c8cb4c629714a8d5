"""
ALE layer transport on the device: the producer of the tracer step's inputs (Qx, Qy, V0, V1, ok) in the ALE
vertical coordinate, as HIP kernels of ``libocx.so``.

The functions keep the names, argument meaning and errors of ``core/vertical/ale.py`` (torch tensors in HBM in
place of NumPy arrays) and ``LayerTransport`` follows ``ALELayerTransport.__call__``
(core/components/layer_transport.py:111-156):

    transport_velocity  (u and v)  ->  ppm_transport_x or upwind_transport_x (u-faces, v-faces)
        ->  open_face_transports (open edges)  ->  limit_outflow
        ->  lagrangian_thickness + layer volumes and wet mask  ->  freshwater_volumes (h_fw, V_fw)

Results are bit-identical to the reference's FP64 arithmetic.  Two calls synchronise with the host, as they must
to behave as the reference does: ``limit_outflow`` (whether to limit, how many passes, its warning or error) and
``lagrangian_thickness`` (the emptied-layer error).

Layout: the reference reaches the v-faces by calling the x functions on swapped axes (``a.swapaxes(1, 2)``,
``a.T``).  The mirrors here accept exactly that - transposed views - and run the v-face kernels on the underlying
natural (n, ny+1, nx) arrays, returning a transposed view again; ``ppm_transport`` and ``LayerTransport`` take the
natural layouts directly.
"""
from __future__ import annotations

import ctypes
import warnings
from dataclasses import dataclass

import torch

from . import _native
from ._native import check, lib
from .ops import Workspace, _ptr, _stream, periodic_pad

H_PPM_MIN = 2e-10        # positive-definite floor of the PPM thickness [m] (core/vertical/ale.py:298)
DZ_FLOOR = 1e-6          # smallest layer thickness an evaporation deficit leaves [m] (core/vertical/ale.py:63)
_F64 = torch.float64

_WS_PPM = Workspace()
_WS_OUT = Workspace()
_WS_LAG = Workspace()


def _f64(t, shape, name):
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)}")
    if t.dtype != _F64:
        raise ValueError(f"{name}: dtype {t.dtype}, expected float64")
    if not t.is_cuda:
        raise ValueError(f"{name}: must be a GPU tensor")
    return t.contiguous()


def _mask(t, shape, name):
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)}")
    if t.dtype == torch.bool:
        t = t.contiguous().view(torch.uint8)
    if t.dtype != torch.uint8 or not t.is_cuda:
        raise ValueError(f"{name}: must be a bool or uint8 GPU tensor")
    return t.contiguous()


def _swapped(t) -> bool:
    """A transposed view of a contiguous array, as the reference's v-face calls pass (a.swapaxes(1, 2), a.T)."""
    if t.is_contiguous():
        return False
    return t.transpose(-1, -2).is_contiguous()


def _natural(t):
    """The (..., y, x) array behind a swapped (..., x, y) argument (no copy when it is a view of one)."""
    return t.transpose(-1, -2).contiguous()


# ------------------------------------------------------------------------------------------------ transport velocity
def transport_velocity(u_new, u_old, last_wet, w, L, D, theta: float, theta_bc: float | None = None):
    """
    (u^θ, Q_bt) of the ALE step: u^θ_k = θ_bc u_k^{n+1} + (1 - θ_bc) u_k^n + (θ - θ_bc)(ū^{n+1} - ū^n) with u^n = 0
    on faces dry in the last step, Q_bt = L D Σ_k w_k u^θ_k.  u (n, a, b); last_wet, L, D (a, b); w (n, a, b).
    core/vertical/ale.py:811-824 (transport_velocity) / :789-808 (its kernel).
    """
    n, a, b = u_new.shape
    if tuple(u_old.shape) != (n, a, b) or tuple(w.shape) != (n, a, b) or any(
            tuple(x.shape) != (a, b) for x in (last_wet, L, D)):
        raise ValueError(f"transport_velocity: shapes u_new {tuple(u_new.shape)}, u_old {tuple(u_old.shape)}, "
                         f"w {tuple(w.shape)}, last_wet/L/D {tuple(last_wet.shape)}/{tuple(L.shape)}/"
                         f"{tuple(D.shape)}")
    un, uo, ww = (_f64(x, (n, a, b), nm) for x, nm in ((u_new, "u_new"), (u_old, "u_old"), (w, "w")))
    Lc, Dc = _f64(L, (a, b), "L"), _f64(D, (a, b), "D")
    wet = _mask(last_wet, (a, b), "last_wet")
    ut = torch.empty_like(un)
    Qb = torch.empty((a, b), dtype=_F64, device=un.device)
    tb = theta if theta_bc is None else theta_bc
    check(lib().ocx_transport_velocity_f64(_ptr(un), _ptr(uo), _ptr(wet), _ptr(ww), _ptr(Lc), _ptr(Dc), n, a * b,
                                           float(theta), float(tb), _ptr(ut), _ptr(Qb), _stream()),
          "ocx_transport_velocity_f64")
    return ut, Qb


# ------------------------------------------------------------------------------------------------ transport thickness
def face_thickness_axis(h, H, active, clip: bool, axis: int, periodic: tuple[bool, bool] = (False, False)):
    """(q_pos, q_neg) at the faces of ``axis`` (0: u-faces (n, ny, nx+1), 1: v-faces (n, ny+1, nx)) from the
    natural-layout h (n, ny, nx), H, active (ny, nx).  On periodic axes the inputs are padded by one wrapped cell
    and the faces cropped, as ``ALECoordinate.transport_thickness`` does (core/components/vertical_coordinate.py:
    125-140)."""
    px, py = periodic
    if px or py:
        act = _mask(active, tuple(h.shape[1:]), "active")
        qp, qn = face_thickness_axis(periodic_pad(h, "c", 1, px, py), periodic_pad(H, "c", 1, px, py) if clip else None,
                                     periodic_pad(act, "c", 1, px, py), clip, axis)
        k = "u" if axis == 0 else "v"
        return (periodic_pad(qp, k, 1, px, py, crop=True), periodic_pad(qn, k, 1, px, py, crop=True))
    n, ny, nx = h.shape
    hc = _f64(h, (n, ny, nx), "h")
    act = _mask(active, (ny, nx), "active")
    Hc = _f64(H, (ny, nx), "H") if clip else None
    shape = (n, ny, nx + 1) if axis == 0 else (n, ny + 1, nx)
    qp = torch.empty(shape, dtype=_F64, device=hc.device)
    qn = torch.empty_like(qp)
    check(lib().ocx_face_transport_thickness_f64(_ptr(hc), _ptr(Hc), _ptr(act), n, ny, nx, int(axis), int(bool(clip)),
                                                 _ptr(qp), _ptr(qn), _stream()), "ocx_face_transport_thickness_f64")
    return qp, qn


def face_transport_thickness(h, H, active, clip: bool):
    """
    Transport thickness at the u-faces (q_e, q_w), each (n, ny, nx+1): the donor layer (NaN read as 0), cut at the
    shallower bed with ``clip``; 0 where a side is land.  Called on swapped axes (h.swapaxes(1, 2), H.T, active.T)
    it gives the v-face (q_n, q_s) as (n, nx, ny+1) views.  core/vertical/ale.py:212-230 (kernel :168-199).
    """
    if _swapped(h):
        qp, qn = face_thickness_axis(_natural(h), _natural(H) if clip else None, _natural(active), clip, 1)
        return qp.transpose(1, 2), qn.transpose(1, 2)
    return face_thickness_axis(h, H, active, clip, 0)


# ------------------------------------------------------------------------------------------------ PPM
def ppm_edges_axis(h, active, axis: int, h_min: float = H_PPM_MIN):
    """Edge thicknesses (low, high) along ``axis`` of the natural-layout h (n, ny, nx)."""
    n, ny, nx = h.shape
    hc = _f64(h, (n, ny, nx), "h")
    act = _mask(active, (ny, nx), "active")
    lo, hi = torch.empty_like(hc), torch.empty_like(hc)
    check(lib().ocx_ppm_edges_f64(_ptr(hc), _ptr(act), n, ny, nx, int(axis), float(h_min), _ptr(lo), _ptr(hi),
                                  _stream()), "ocx_ppm_edges_f64")
    return lo, hi


def ppm_edges_x(h, active, h_min: float = H_PPM_MIN):
    """West and east edge thickness of every cell (core/vertical/ale.py:440-476, kernel :578-623); on swapped axes
    the south and north edges."""
    if _swapped(h):
        lo, hi = ppm_edges_axis(_natural(h), _natural(active), 1, h_min)
        return lo.transpose(1, 2), hi.transpose(1, 2)
    return ppm_edges_axis(h, active, 0, h_min)


def ppm_transport(h, ut, L, Q_bt, q_pos, q_neg, active, dt: float, d_t, axis: int, max_iter: int = 20, out=None,
                  workspace: Workspace | None = None, return_iterations: bool = False):
    """
    Layer volume fluxes through the faces of ``axis`` with PPM thickness, the uniform Newton correction per face
    column and the residual share, in natural layouts: h (n, ny, nx); ut, q_pos, q_neg, out (n, ny, nx+1) for
    axis 0 or (n, ny+1, nx) for axis 1; L, Q_bt one layer of faces; d_t the cell width along the axis (ny, nx).
    ``return_iterations`` also returns the Newton passes taken (a host synchronisation).
    core/vertical/ale.py:743-764 (ppm_transport_x), kernels :578-740.
    """
    n, ny, nx = h.shape
    fshape = (n, ny, nx + 1) if axis == 0 else (n, ny + 1, nx)
    hc = _f64(h, (n, ny, nx), "h")
    u = _f64(ut, fshape, "ut")
    Lc, Qb = _f64(L, fshape[1:], "L"), _f64(Q_bt, fshape[1:], "Q_bt")
    qp, qn = _f64(q_pos, fshape, "q_pos"), _f64(q_neg, fshape, "q_neg")
    act = _mask(active, (ny, nx), "active")
    dd = _f64(d_t, (ny, nx), "d_t")
    o = torch.empty(fshape, dtype=_F64, device=hc.device) if out is None else out
    if tuple(o.shape) != fshape or o.dtype != _F64 or not o.is_contiguous():
        raise ValueError(f"out: a contiguous float64 tensor of shape {fshape}")
    Lb = lib()
    ws = (workspace or _WS_PPM).get(int(Lb.ocx_ppm_transport_workspace_bytes(n, ny, nx, int(axis), int(max_iter))))
    check(Lb.ocx_ppm_transport_f64(_ptr(hc), _ptr(u), _ptr(Lc), _ptr(Qb), _ptr(qp), _ptr(qn), _ptr(act), _ptr(dd),
                                   n, ny, nx, int(axis), float(dt), int(max_iter), _ptr(o), _ptr(ws), ws.numel(),
                                   _stream()), "ocx_ppm_transport_f64")
    if return_iterations:
        it = Lb.ocx_ppm_transport_iterations(_ptr(ws), n, ny, nx, int(axis), int(max_iter), _stream())
        if it < 0:
            check(it, "ocx_ppm_transport_iterations")
        return o, it
    return o


def ppm_transport_x(h, ut, L, Q_bt, q_e, q_w, active, dt: float, dx_t, max_iter: int = 20):
    """
    Layer volume fluxes through the u-faces with PPM thickness (the reference's signature and result,
    core/vertical/ale.py:743-764).  Called as the reference calls it for the v-faces - every array a swapped view
    (h.swapaxes(1, 2), ut.swapaxes(1, 2), Lv.T, Qby.T, act.T, dy_t.T) and q_n, q_s in the (n, nx, ny+1) orientation -
    it runs the v-face kernels and returns the (n, nx, ny+1) view of the natural result.
    """
    if _swapped(h):
        out = ppm_transport(_natural(h), _natural(ut), _natural(L), _natural(Q_bt), _natural(q_e), _natural(q_w),
                            _natural(active), dt, _natural(dx_t), 1, max_iter)
        return out.transpose(1, 2)
    return ppm_transport(h, ut, L, Q_bt, q_e, q_w, active, dt, dx_t, 0, max_iter)


def upwind_transport(ut, L, Q_bt, q_pos, q_neg, axis: int, out=None):
    """
    Layer volume fluxes through the faces of ``axis`` with the donor-cell thickness, corrected to the free-surface
    flux Q_bt (the ``thickness_flux="upwind"`` option), in natural layouts: ut, q_pos, q_neg, out (n, ny, nx+1) for
    axis 0 or (n, ny+1, nx) for axis 1; L, Q_bt one layer of faces.  core/vertical/ale.py:233-251.
    """
    n = ut.shape[0]
    fy, fx = ut.shape[1:]
    ny, nx = (fy, fx - 1) if axis == 0 else (fy - 1, fx)
    fshape = (n, fy, fx)
    u = _f64(ut, fshape, "ut")
    Lc, Qb = _f64(L, fshape[1:], "L"), _f64(Q_bt, fshape[1:], "Q_bt")
    qp, qn = _f64(q_pos, fshape, "q_pos"), _f64(q_neg, fshape, "q_neg")
    o = torch.empty(fshape, dtype=_F64, device=u.device) if out is None else out
    if tuple(o.shape) != fshape or o.dtype != _F64 or not o.is_contiguous():
        raise ValueError(f"out: a contiguous float64 tensor of shape {fshape}")
    check(lib().ocx_upwind_transport_f64(_ptr(u), _ptr(Lc), _ptr(Qb), _ptr(qp), _ptr(qn), n, ny, nx, int(axis),
                                         _ptr(o), _stream()), "ocx_upwind_transport_f64")
    return o


def upwind_transport_x(q_e, q_w, ut, L, Q_bt):
    """
    Layer volume fluxes through the u-faces with the donor-cell thickness (the reference's signature and result,
    core/vertical/ale.py:233-251).  Called as the reference calls it for the v-faces - q_n, q_s, ut in the
    (n, nx, ny+1) orientation, Lv.T, Qby.T - it runs on the natural v-face arrays and returns the swapped view.
    """
    if _swapped(ut):
        out = upwind_transport(_natural(ut), _natural(L), _natural(Q_bt), _natural(q_e), _natural(q_w), 1)
        return out.transpose(1, 2)
    return upwind_transport(ut, L, Q_bt, q_e, q_w, 0)


_EDGE_BITS = {"west": 1, "east": 2, "south": 4, "north": 8}


def open_face_transports(Qx, Qy, h, un, vn, u_new, v_new, open_edges, open_Du, open_Qu, open_Dv, open_Qv, Lu, Lv,
                         theta: float, theta_bc: float | None = None):
    """
    In place on Qx (n, ny, nx+1) and Qy (n, ny+1, nx): on the open edges (names 'west', 'east', 'south', 'north'),
    every face whose free-surface face depth open_D > 0 takes the layer fluxes of the interior cell's thicknesses
    adding up to the Flather flux open_Q (``OpenBoundary.open_face_transports``, core/components/open_boundary.py:
    334-351; ``open_face_transport``, core/vertical/ale.py:841-861); un, vn the old velocities.  Returns (Qx, Qy).
    """
    n, ny, nx = h.shape
    mask = 0
    for e in open_edges:
        if e not in _EDGE_BITS:
            raise ValueError(f"unknown open edge {e!r}")
        mask |= _EDGE_BITS[e]
    if not mask:
        return Qx, Qy
    for t, nm in ((Qx, "Qx"), (Qy, "Qy")):
        if t.dtype != _F64 or not t.is_cuda or not t.is_contiguous():
            raise ValueError(f"{nm}: a contiguous float64 GPU tensor (written in place)")
    _f64(Qx, (n, ny, nx + 1), "Qx")
    _f64(Qy, (n, ny + 1, nx), "Qy")
    hc = _f64(h, (n, ny, nx), "h")
    uo, vo = _f64(un, (n, ny, nx + 1), "un"), _f64(vn, (n, ny + 1, nx), "vn")
    u1, v1 = _f64(u_new, (n, ny, nx + 1), "u_new"), _f64(v_new, (n, ny + 1, nx), "v_new")
    Du, Qu = _f64(open_Du, (ny, nx + 1), "open_Du"), _f64(open_Qu, (ny, nx + 1), "open_Qu")
    Dv, Qv = _f64(open_Dv, (ny + 1, nx), "open_Dv"), _f64(open_Qv, (ny + 1, nx), "open_Qv")
    Luc, Lvc = _f64(Lu, (ny, nx + 1), "Lu"), _f64(Lv, (ny + 1, nx), "Lv")
    tb = float(theta if theta_bc is None else theta_bc)
    check(lib().ocx_open_face_transports_f64(_ptr(Qx), _ptr(Qy), _ptr(hc), _ptr(uo), _ptr(vo), _ptr(u1), _ptr(v1),
                                             _ptr(Du), _ptr(Qu), _ptr(Dv), _ptr(Qv), _ptr(Luc), _ptr(Lvc), n, ny, nx,
                                             mask, tb, _stream()), "ocx_open_face_transports_f64")
    return Qx, Qy


# ------------------------------------------------------------------------------------------------ outflow limiter
def _outflow_call(h, Qx, Qy, area, dt, active, h_floor, max_pass, periodic, check_only):
    n, ny, nx = h.shape
    Lb = lib()
    ws = _WS_OUT.get(int(Lb.ocx_limit_outflow_workspace_bytes(n, ny, nx)))
    nl, st = ctypes.c_longlong(0), ctypes.c_int(0)
    where = (ctypes.c_longlong * 3)()
    val = ctypes.c_double(0.0)
    rc = Lb.ocx_limit_outflow_f64(_ptr(h), _ptr(Qx), _ptr(Qy), _ptr(area), _ptr(active), n, ny, nx, float(dt),
                                  float(h_floor), int(max_pass), int(bool(periodic[0])), int(bool(periodic[1])),
                                  int(bool(check_only)), _ptr(ws), ws.numel(), ctypes.byref(nl), ctypes.byref(st),
                                  where, ctypes.byref(val), _stream())
    check(rc, "limit_outflow")
    return nl.value, st.value, tuple(where), val.value


def limit_outflow(h, Qx, Qy, area, dt: float, active, h_floor: float = H_PPM_MIN, max_pass: int = 50,
                  periodic: tuple[bool, bool] = (False, False)):
    """
    Outflow limiter of the Lagrangian step: no layer of a cell sends out more than (h - h_floor)+ A in Δt; each
    face's cut goes back to its other layers so Σ_k Q_k stays the free-surface flux.  Returns (Qx, Qy, n_limited):
    the inputs themselves when nothing exceeds, else new tensors.  Warns (RuntimeWarning) when some layer still
    exceeds after ``max_pass`` but h stays >= -h_floor; raises RuntimeError where the reference raises.
    core/vertical/ale.py:331-437.  Host-synchronous.
    """
    n, ny, nx = h.shape
    hc = _f64(h, (n, ny, nx), "h")
    qx = _f64(Qx, (n, ny, nx + 1), "Qx")
    qy = _f64(Qy, (n, ny + 1, nx), "Qy")
    ar = _f64(area, (ny, nx), "area")
    act = _mask(active, (ny, nx), "active")
    _, st, _, _ = _outflow_call(hc, qx, qy, ar, dt, act, h_floor, max_pass, periodic, True)
    if st == _native.OUTFLOW_NOOP:
        return Qx, Qy, 0
    qx, qy = qx.clone(), qy.clone()
    nl, st, where, val = _outflow_call(hc, qx, qy, ar, dt, act, h_floor, max_pass, periodic, False)
    if st == _native.OUTFLOW_ACCEPTED:
        k, j, i = where
        warnings.warn(f"limit_outflow: some layer still sends out more than it holds after {max_pass} passes, but "
                      f"its inflow keeps h >= 0 (min {val:.3g} m at layer {k}, cell j={j}, i={i}); the FCT bound "
                      f"is weaker there", RuntimeWarning, stacklevel=2)
    return qx, qy, nl


# ------------------------------------------------------------------------------------------------ Lagrangian step
def lagrangian_thickness(h, divq, dt: float, area, tol: float = 1e-9):
    """
    h† = h - Δt (∇·Q_k) / A, values in [-tol, 0) set to 0; a layer emptied beyond that raises RuntimeError
    naming it.  h, divq (n, ny, nx); area (ny, nx).  core/vertical/ale.py:278-295.  Host-synchronous.
    """
    n, ny, nx = h.shape
    hc = _f64(h, (n, ny, nx), "h")
    dv = _f64(divq, (n, ny, nx), "divq")
    ar = _f64(area, (ny, nx), "area")
    out = torch.empty_like(hc)
    Lb = lib()
    ws = _WS_LAG.get(int(Lb.ocx_lagrangian_workspace_bytes()))
    bad = ctypes.c_longlong(-1)
    check(Lb.ocx_lagrangian_thickness_f64(_ptr(hc), _ptr(dv), None, None, _ptr(ar), None, n, ny, nx, float(dt),
                                          float(tol), _ptr(out), None, None, None, _ptr(ws), ws.numel(),
                                          ctypes.byref(bad), _stream()), "lagrangian_thickness")
    return out


# ------------------------------------------------------------------------------------------------ the producer
@dataclass
class LayerFluxes:
    """The tracer step's inputs from the ALE layer transport (core/components/layer_transport.py:29-41)."""
    Qx: torch.Tensor          # (n, ny, nx+1) [m³ s⁻¹]
    Qy: torch.Tensor          # (n, ny+1, nx)
    V0: torch.Tensor          # layer volumes before [m³]
    V1: torch.Tensor          # after (h† A)
    ok: torch.Tensor          # wet cell-layers after the step (uint8)
    h_dagger: torch.Tensor    # Lagrangian thickness [m]
    n_limited: int = 0
    V_fw: torch.Tensor | None = None    # with freshwater: h_fw A [m³]
    h_fw: torch.Tensor | None = None    # h† with dt fw on the top layer [m]


class LayerTransport:
    """
    The ALE layer-transport step on the device (``ALELayerTransport.__call__``, core/components/layer_transport.py:
    111-156): per call the transport velocities of both face directions, the PPM
    fluxes with their Newton correction (or, with ``thickness_flux="upwind"``, the donor-cell fluxes), the outflow
    limiter, the Lagrangian thickness and the layer volumes.  Static geometry is given once: Lu (ny, nx+1), Lv (ny+1, nx) face lengths; area, active,
    dx_t, dy_t (ny, nx); dt; theta (free-surface time weight).
    """

    def __init__(self, *, Lu, Lv, area, active, dx_t, dy_t, dt: float, theta: float, max_iter: int = 20,
                 h_floor: float = H_PPM_MIN, max_pass: int = 50, periodic: tuple[bool, bool] = (False, False),
                 thickness_flux: str = "ppm"):
        if thickness_flux not in ("ppm", "upwind"):
            raise ValueError(f"thickness_flux {thickness_flux!r}: 'ppm' or 'upwind'")
        self.thickness_flux = thickness_flux
        ny, nx = area.shape
        self.Lu, self.Lv = _f64(Lu, (ny, nx + 1), "Lu"), _f64(Lv, (ny + 1, nx), "Lv")
        self.area, self.dx_t, self.dy_t = (_f64(x, (ny, nx), nm) for x, nm in ((area, "area"), (dx_t, "dx_t"),
                                                                                (dy_t, "dy_t")))
        self.active = _mask(active, (ny, nx), "active")
        self.dt, self.theta, self.max_iter = float(dt), float(theta), int(max_iter)
        self.h_floor, self.max_pass = float(h_floor), int(max_pass)
        self.periodic = (bool(periodic[0]), bool(periodic[1]))
        self.ws_x, self.ws_y = Workspace(), Workspace()
        self.iterations = (0, 0)

    def __call__(self, h, u_new, u_old, v_new, v_old, wet_u, wet_v, wu, wv, Du, Dv, qe, qw, qn, qs,
                 theta_bc: float | None = None, count_iterations: bool = False, fw=None, open_=(),
                 open_Du=None, open_Qu=None, open_Dv=None, open_Qv=None) -> LayerFluxes:
        """h (n, ny, nx) layer thickness at step n; u (n, ny, nx+1), v (n, ny+1, nx) new and old velocities;
        wet_u / wet_v faces wet in the last free-surface step; wu, wv face fractions; Du, Dv face depths;
        qe, qw (u-faces) and qn, qs (v-faces, natural (n, ny+1, nx) layout) transport thicknesses."""
        utx, Qbx = transport_velocity(u_new, u_old, wet_u, wu, self.Lu, Du, self.theta, theta_bc)
        uty, Qby = transport_velocity(v_new, v_old, wet_v, wv, self.Lv, Dv, self.theta, theta_bc)
        px, py = self.periodic
        if self.thickness_flux == "upwind":
            # donor-cell thickness, face by face: no padding (layer_transport.py:131-133)
            Qx = upwind_transport(utx, self.Lu, Qbx, qe, qw, 0)
            Qy = upwind_transport(uty, self.Lv, Qby, qn, qs, 1)
            if count_iterations:
                Qx, Qy = (Qx, 0), (Qy, 0)
        elif px or py:
            # the PPM stencils on inputs padded by 4 wrapped cells, the fluxes cropped (layer_transport.py:122-129)
            def P(a, kind):
                return periodic_pad(a, kind, 4, px, py)
            hp, ap = P(h, "c"), P(self.active, "c")
            Qx = ppm_transport(hp, P(utx, "u"), P(self.Lu, "u"), P(Qbx, "u"), P(qe, "u"), P(qw, "u"), ap, self.dt,
                               P(self.dx_t, "c"), 0, self.max_iter, workspace=self.ws_x,
                               return_iterations=count_iterations)
            Qy = ppm_transport(hp, P(uty, "v"), P(self.Lv, "v"), P(Qby, "v"), P(qn, "v"), P(qs, "v"), ap, self.dt,
                               P(self.dy_t, "c"), 1, self.max_iter, workspace=self.ws_y,
                               return_iterations=count_iterations)
        else:
            Qx = ppm_transport(h, utx, self.Lu, Qbx, qe, qw, self.active, self.dt, self.dx_t, 0, self.max_iter,
                               workspace=self.ws_x, return_iterations=count_iterations)
            Qy = ppm_transport(h, uty, self.Lv, Qby, qn, qs, self.active, self.dt, self.dy_t, 1, self.max_iter,
                               workspace=self.ws_y, return_iterations=count_iterations)
        if count_iterations:
            (Qx, ix), (Qy, iy) = Qx, Qy
            self.iterations = (ix, iy)
        if (px or py) and self.thickness_flux == "ppm":
            Qx = periodic_pad(Qx, "u", 4, px, py, crop=True)
            Qy = periodic_pad(Qy, "v", 4, px, py, crop=True)
        if open_:
            # the open edges' faces from the interior cells and the Flather flux (layer_transport.py:134-135)
            Qx, Qy = Qx.contiguous(), Qy.contiguous()
            open_face_transports(Qx, Qy, h, u_old, v_old, u_new, v_new, open_, open_Du, open_Qu, open_Dv, open_Qv,
                                 self.Lu, self.Lv, self.theta, theta_bc)
        Qx, Qy, n_lim = limit_outflow(h, Qx, Qy, self.area, self.dt, self.active, self.h_floor, self.max_pass,
                                      periodic=self.periodic)
        n, ny, nx = h.shape
        h = _f64(h, (n,) + tuple(self.area.shape), "h")
        h_lag = torch.empty_like(h)
        V0, V1 = torch.empty_like(h), torch.empty_like(h)
        ok = torch.empty((n, ny, nx), dtype=torch.uint8, device=h.device)
        Lb = lib()
        ws = _WS_LAG.get(int(Lb.ocx_lagrangian_workspace_bytes()))
        bad = ctypes.c_longlong(-1)
        Qx, Qy = Qx.contiguous(), Qy.contiguous()
        check(Lb.ocx_lagrangian_thickness_f64(_ptr(h), None, _ptr(Qx), _ptr(Qy),
                                              _ptr(self.area), _ptr(self.active), n, ny, nx, self.dt, 1e-9,
                                              _ptr(h_lag), _ptr(V0), _ptr(V1), _ptr(ok), _ptr(ws), ws.numel(),
                                              ctypes.byref(bad), _stream()), "lagrangian_thickness")
        h_fw = V_fw = None
        if fw is not None:
            # real freshwater: the top layer of h† gains dt fw; a deficit below the floor comes from the layers
            # below (layer_transport.py:148-155); h† and V1 stay without it
            from .freshwater import freshwater_volumes
            h_fw, V_fw, _ = freshwater_volumes(h_lag, fw, DZ_FLOOR, scale=self.dt, area=self.area)
        return LayerFluxes(Qx, Qy, V0, V1, ok, h_lag, n_lim, V_fw, h_fw)
