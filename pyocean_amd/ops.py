"""
Device-resident entry points: torch tensors in HBM in and out, the HIP kernels of ``libocx.so`` underneath.

torch is used only as plumbing: it owns the device memory (its caching allocator) and the current stream.  Every
function checks shapes on the host and raises ValueError before any launch, so a kernel never sees a shape it does
not expect.  Tensors must be FP64 (``ok``: uint8 or bool), C-contiguous, on the current GPU.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native
from ._native import check, lib

_F64 = torch.float64


def require_gpu() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("pyocean_amd runs on an AMD GPU (ROCm); no GPU is visible")
    return torch.device("cuda", torch.cuda.current_device())


def _ptr(t):
    """The device address of ``t`` as a ctypes argument that keeps ``t`` alive while the call runs (a temporary
    tensor in an argument list would otherwise be freed, and its memory reused, before the kernel reads it)."""
    if t is None:
        return None
    p = ctypes.c_void_p(t.data_ptr())
    p._tensor = t
    return p


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _need(t, shape, name, dtype=_F64):
    if t is None:
        raise ValueError(f"{name} is None")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)}")
    if t.dtype != dtype:
        raise ValueError(f"{name}: dtype {t.dtype}, expected {dtype}")
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError(f"{name}: must be a contiguous GPU tensor")
    return t


def _ok(ok, shape):
    if ok.dtype == torch.bool:
        ok = ok.view(torch.uint8)
    return _need(ok, shape, "ok", torch.uint8)


class Workspace:
    """A growing device scratch buffer, one per stream, reused across the calls made on that stream.  The caching
    allocator hands a freed block only to work ordered after the free on the same stream, so growing the buffer of
    one stream never releases memory that another stream's kernels may still be using, and calls on two streams
    never share a buffer."""

    def __init__(self):
        self.bufs = {}

    def get(self, nbytes: int) -> torch.Tensor:
        s = torch.cuda.current_stream(require_gpu())
        key = (s.device_index, s.cuda_stream)
        buf = self.bufs.get(key)
        if buf is None or buf.numel() < nbytes:
            self.bufs.pop(key, None)
            buf = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=s.device)
            self.bufs[key] = buf
        return buf


_WS = Workspace()
_WS_CONTENT = Workspace()


def flags_of(open_mask: int = 0, js: bool = False, unlimited: bool = False, up: int = 0,
             centred: bool = False) -> int:
    """Flag word of the C-ABI: open edges, WENO5-JS faces, the unlimited FCT result, ``up`` = 3 / 1 for the
    UP3 / donor-cell faces of the RK3 step, or the centred faces of the RHS and faces_h."""
    if up not in (0, 1, 3):
        raise ValueError(f"up = {up}: 0, 1 or 3")
    return (int(open_mask) | (_native.WENO_JS if js else 0) | (_native.UNLIMITED if unlimited else 0)
            | {0: 0, 1: _native.UP1, 3: _native.UP3}[up] | (_native.CENTRED if centred else 0))


PRECISIONS = ("f64", "split")


def advect_fct(C, V0, V1, Qx, Qy, ok, dt, inflow=None, flags=0, flux_face_i=None, open_flux=False, out=None,
               workspace: Workspace | None = None, precision: str = "f64"):
    """
    One SSP-RK3 + FCT step on device tensors: ``precision`` "f64" (``ocx_advect_fct_f64``, bit-identical to the
    reference) or "split" (``ocx_advect_fct_split``: FP32 reconstruction, FP64 accumulation and limiter).  Returns
    ``out`` or (out, fx_i, open_flux) when ``flux_face_i`` / ``open_flux`` are asked for (None where not).
    """
    if precision not in PRECISIONS:
        raise ValueError(f"precision {precision!r}: use one of {PRECISIONS}")
    nt, n, ny, nx = C.shape
    _need(C, (nt, n, ny, nx), "C")
    for t, s, nm in ((V0, (n, ny, nx), "V0"), (V1, (n, ny, nx), "V1"), (Qx, (n, ny, nx + 1), "Qx"),
                     (Qy, (n, ny + 1, nx), "Qy")):
        _need(t, s, nm)
    ok = _ok(ok, (n, ny, nx))
    if inflow is not None:
        _need(inflow, C.shape, "inflow")
    out = torch.empty_like(C) if out is None else _need(out, C.shape, "out")
    fxi = torch.empty((nt, n, ny), dtype=_F64, device=C.device) if flux_face_i is not None else None
    of = torch.empty((nt,), dtype=_F64, device=C.device) if open_flux else None
    L = lib()
    split = precision == "split"
    nb = (L.ocx_fct_split_workspace_bytes if split else L.ocx_fct_workspace_bytes)(nt, n, ny, nx)
    ws = (workspace or _WS).get(nb)
    fn = L.ocx_advect_fct_split if split else L.ocx_advect_fct_f64
    rc = fn(_ptr(C), _ptr(V0), _ptr(V1), _ptr(Qx), _ptr(Qy), _ptr(ok), _ptr(inflow), nt, n, ny, nx, float(dt),
            int(flags), -1 if flux_face_i is None else int(flux_face_i), _ptr(out), _ptr(fxi), _ptr(of), _ptr(ws),
            ws.numel(), _stream())
    check(rc, "ocx_advect_fct_split" if split else "ocx_advect_fct_f64")
    if flux_face_i is None and not open_flux:
        return out
    return out, fxi, of


def _up_faces(flags, Fz) -> bool:
    """True for the UP3 / UP1 / centred faces, whose RHS always forms the z faces and so needs Fz."""
    up = bool(int(flags) & (_native.UP3 | _native.UP1 | _native.CENTRED))
    if up and Fz is None:
        raise ValueError("the upwind3 / upwind1 / centred faces need Fz (their RHS always forms the z faces)")
    return up


def advect_rk3(C, V0, V1, Qx, Qy, Fz, ok, dt, inflow=None, flags=0, vert=None, flux_face_i=None, out=None,
               workspace: Workspace | None = None):
    """One SSP-RK3 step (x, y and z sweeps) on device tensors (``ocx_advect_rk3_f64``)."""
    nt, n, ny, nx = C.shape
    _need(C, (nt, n, ny, nx), "C")
    for t, s, nm in ((V0, (n, ny, nx), "V0"), (V1, (n, ny, nx), "V1"), (Qx, (n, ny, nx + 1), "Qx"),
                     (Qy, (n, ny + 1, nx), "Qy")):
        _need(t, s, nm)
    if Fz is not None:
        _need(Fz, (n + 1, ny, nx), "Fz")
    up = _up_faces(flags, Fz)
    if vert is None:
        vert = not up and Fz is not None and bool(torch.any(Fz != 0.0).item())
    ok = _ok(ok, (n, ny, nx))
    if inflow is not None:
        _need(inflow, C.shape, "inflow")
    out = torch.empty_like(C) if out is None else _need(out, C.shape, "out")
    fxi = torch.empty((nt, n, ny), dtype=_F64, device=C.device) if flux_face_i is not None else None
    L = lib()
    nb = L.ocx_rk3_workspace_bytes(nt, n, ny, nx)
    ws = (workspace or _WS).get(nb)
    fz = Fz if vert or up else None
    rc = L.ocx_advect_rk3_f64(_ptr(C), _ptr(V0), _ptr(V1), _ptr(Qx), _ptr(Qy), _ptr(fz), _ptr(ok),
                              _ptr(inflow), nt, n, ny, nx, float(dt), int(flags), int(bool(vert)),
                              -1 if flux_face_i is None else int(flux_face_i), _ptr(out), _ptr(fxi), _ptr(ws),
                              ws.numel(), _stream())
    check(rc, "ocx_advect_rk3_f64")
    return out if flux_face_i is None else (out, fxi)


def advect_pc(C, V0, V1, Qx, Qy, Fz, ok, dt, inflow=None, flags=0, vert=None, flux_face_i=None, out=None,
              workspace: Workspace | None = None):
    """One 'pc' step on device tensors (``ocx_advect_pc_f64``): the centred half step, then the full step with the
    faces of ``flags`` (WENO5-Z, WENO5-JS, UP3 or UP1) of the mid-step tracer.  Returns out or (out, fx_i)."""
    nt, n, ny, nx = C.shape
    _need(C, (nt, n, ny, nx), "C")
    for t, s, nm in ((V0, (n, ny, nx), "V0"), (V1, (n, ny, nx), "V1"), (Qx, (n, ny, nx + 1), "Qx"),
                     (Qy, (n, ny + 1, nx), "Qy")):
        _need(t, s, nm)
    if Fz is None:
        raise ValueError("the pc step needs Fz (its centred predictor always forms the z faces)")
    _need(Fz, (n + 1, ny, nx), "Fz")
    up = _up_faces(flags, Fz)
    if vert is None:
        vert = not up and bool(torch.any(Fz != 0.0).item())
    ok = _ok(ok, (n, ny, nx))
    if inflow is not None:
        _need(inflow, C.shape, "inflow")
    out = torch.empty_like(C) if out is None else _need(out, C.shape, "out")
    fxi = torch.empty((nt, n, ny), dtype=_F64, device=C.device) if flux_face_i is not None else None
    L = lib()
    ws = (workspace or _WS).get(L.ocx_rk3_workspace_bytes(nt, n, ny, nx))
    rc = L.ocx_advect_pc_f64(_ptr(C), _ptr(V0), _ptr(V1), _ptr(Qx), _ptr(Qy), _ptr(Fz), _ptr(ok), _ptr(inflow), nt, n,
                             ny, nx, float(dt), int(flags), int(bool(vert)),
                             -1 if flux_face_i is None else int(flux_face_i), _ptr(out), _ptr(fxi), _ptr(ws),
                             ws.numel(), _stream())
    check(rc, "ocx_advect_pc_f64")
    return out if flux_face_i is None else (out, fxi)


def faces_h(c, Qx, Qy, ok, ext=None, flags=0):
    """Horizontal face fluxes (fx, fy) of the stack c (``ocx_faces_h_f64``): WENO, or UP3 / UP1 / centred by the
    flags."""
    nt, n, ny, nx = c.shape
    _need(c, (nt, n, ny, nx), "c")
    _need(Qx, (n, ny, nx + 1), "Qx")
    _need(Qy, (n, ny + 1, nx), "Qy")
    ok = _ok(ok, (n, ny, nx))
    if ext is not None:
        _need(ext, c.shape, "ext")
    fx = torch.empty((nt, n, ny, nx + 1), dtype=_F64, device=c.device)
    fy = torch.empty((nt, n, ny + 1, nx), dtype=_F64, device=c.device)
    check(lib().ocx_faces_h_f64(_ptr(c), _ptr(Qx), _ptr(Qy), _ptr(ok), _ptr(ext), nt, n, ny, nx, int(flags), _ptr(fx),
                                _ptr(fy), _stream()), "ocx_faces_h_f64")
    return fx, fy


def faces(c, flux, ok, axis: int, flags: int = 0):
    """
    Face values (not fluxes) of c along ``axis`` (``ocx_faces_f64``): WENO5-Z, or by the flags WENO5-JS, UP3, UP1
    or centred, upwind-ordered by the sign of ``flux`` (which also sets the number of faces; unused by the centred
    faces).  c, flux and ok broadcast against each other on every other axis, as the reference's face functions
    do (core/dynamics/sigma_baroclinic.py:564-578).  Returns the (lead..., nf) array with the face axis at ``axis``.
    """
    cm = torch.movedim(c, axis, -1)
    fm = torch.movedim(flux, axis, -1)
    om = torch.movedim(ok.view(torch.uint8) if ok.dtype == torch.bool else ok, axis, -1)
    for t, nm in ((cm, "c"), (fm, "flux")):
        if t.dtype != _F64 or not t.is_cuda:
            raise ValueError(f"{nm}: a float64 GPU tensor")
    if om.dtype != torch.uint8 or not om.is_cuda:
        raise ValueError("ok: a bool or uint8 GPU tensor")
    lead = torch.broadcast_shapes(cm.shape[:-1], fm.shape[:-1], om.shape[:-1])
    n, nf = cm.shape[-1], fm.shape[-1]
    ax = axis % (len(lead) + 1)
    place = lambda t, m: torch.movedim(t.expand(tuple(lead) + (m,)), -1, ax).contiguous()     # noqa: E731
    c3, f3, o3 = place(cm, n), place(fm, nf), place(om, n)
    shape = tuple(f3.shape)
    outer = int(np.prod(shape[:ax], dtype=np.int64))
    inner = int(np.prod(shape[ax + 1:], dtype=np.int64))
    out = torch.empty(shape, dtype=_F64, device=c.device)
    if out.numel() == 0:
        return out
    check(lib().ocx_faces_f64(_ptr(c3), _ptr(f3), _ptr(o3), outer, n, nf, inner, int(flags), _ptr(out), _stream()),
          "ocx_faces_f64")
    return out


def open_faces(face, c, Q, axis: int, open_mask: int, ext=None):
    """Open-edge face values in place on ``face`` (``ocx_open_faces_f64``): axis 0 the x faces (nt, nz, ny, nx+1)
    with Q (nz, ny, nx+1), axis 1 the y faces (nt, nz, ny+1, nx) with Q (nz, ny+1, nx); on each open edge of the
    axis the exterior value ``ext`` (default c) where the flux enters, else the boundary cell."""
    nt, n, ny, nx = c.shape
    _need(c, (nt, n, ny, nx), "c")
    fshape = (nt, n, ny, nx + 1) if axis == 0 else (nt, n, ny + 1, nx)
    _need(face, fshape, "face")
    _need(Q, fshape[1:], "Q")
    if ext is not None:
        _need(ext, c.shape, "ext")
    check(lib().ocx_open_faces_f64(_ptr(face), _ptr(c), _ptr(Q), _ptr(ext), nt, n, ny, nx, int(axis),
                                   int(open_mask) & 15, _stream()), "ocx_open_faces_f64")
    return face


def upwind_faces_h(c, Qx, Qy, ext=None, flags=0):
    """Donor-cell face fluxes (``ocx_upwind_faces_h_f64``)."""
    nt, n, ny, nx = c.shape
    _need(c, (nt, n, ny, nx), "c")
    _need(Qx, (n, ny, nx + 1), "Qx")
    _need(Qy, (n, ny + 1, nx), "Qy")
    if ext is not None:
        _need(ext, c.shape, "ext")
    fx = torch.empty((nt, n, ny, nx + 1), dtype=_F64, device=c.device)
    fy = torch.empty((nt, n, ny + 1, nx), dtype=_F64, device=c.device)
    check(lib().ocx_upwind_faces_h_f64(_ptr(c), _ptr(Qx), _ptr(Qy), _ptr(ext), nt, n, ny, nx, int(flags), _ptr(fx),
                                       _ptr(fy), _stream()), "ocx_upwind_faces_h_f64")
    return fx, fy


def rhs(c, Qx, Qy, Fz, ok, ext=None, flags=0, vert=None):
    """Advection RHS -div(Q c) (``ocx_rhs_f64``)."""
    nt, n, ny, nx = c.shape
    _need(c, (nt, n, ny, nx), "c")
    _need(Qx, (n, ny, nx + 1), "Qx")
    _need(Qy, (n, ny + 1, nx), "Qy")
    if Fz is not None:
        _need(Fz, (n + 1, ny, nx), "Fz")
    up = _up_faces(flags, Fz)
    if vert is None:
        vert = not up and Fz is not None and bool(torch.any(Fz != 0.0).item())
    ok = _ok(ok, (n, ny, nx))
    if ext is not None:
        _need(ext, c.shape, "ext")
    out = torch.empty_like(c)
    fz = Fz if vert or up else None
    check(lib().ocx_rhs_f64(_ptr(c), _ptr(Qx), _ptr(Qy), _ptr(fz), _ptr(ok), _ptr(ext), nt, n, ny, nx, int(flags),
                            int(bool(vert)), _ptr(out), _stream()), "ocx_rhs_f64")
    return out


def fct_coefficients(c0, cL, V1, Ax, Ay, ok, dt):
    """Zalesak coefficients (Cx, Cy) (``ocx_fct_coefficients_f64``)."""
    nt, n, ny, nx = c0.shape
    _need(c0, (nt, n, ny, nx), "c0")
    _need(cL, (nt, n, ny, nx), "cL")
    _need(V1, (n, ny, nx), "V1")
    _need(Ax, (nt, n, ny, nx + 1), "Ax")
    _need(Ay, (nt, n, ny + 1, nx), "Ay")
    ok = _ok(ok, (n, ny, nx))
    Cx = torch.empty_like(Ax)
    Cy = torch.empty_like(Ay)
    ws = _WS.get(2 * c0.numel() * 8)
    check(lib().ocx_fct_coefficients_f64(_ptr(c0), _ptr(cL), _ptr(V1), _ptr(Ax), _ptr(Ay), _ptr(ok), float(dt), nt, n,
                                         ny, nx, _ptr(Cx), _ptr(Cy), _ptr(ws), ws.numel(), _stream()),
          "ocx_fct_coefficients_f64")
    return Cx, Cy


def periodic_pad(a, kind: str, halo: int, px: bool, py: bool, crop: bool = False):
    """Wrapped halo on the periodic axes (``ocx_periodic_pad``); ``kind`` 'c', 'u' or 'v'; ``crop``: inverse."""
    if not (px or py):
        return a
    if not a.is_contiguous():
        a = a.contiguous()
    kid = {"c": 0, "u": 1, "v": 2}[kind]
    sy, sx = a.shape[-2], a.shape[-1]
    lead = a.numel() // (sy * sx)
    hy, hx = (halo if py else 0), (halo if px else 0)
    sign = -1 if crop else 1
    dy, dx = sy + 2 * hy * sign, sx + 2 * hx * sign
    if dy < 1 or dx < 1:
        raise ValueError(f"periodic crop of {sy}x{sx} by {hy},{hx}")
    out = torch.empty(a.shape[:-2] + (dy, dx), dtype=a.dtype, device=a.device)
    elem = 8 if a.dtype == _F64 else 1
    if a.dtype not in (_F64, torch.uint8):
        raise ValueError(f"periodic pad of dtype {a.dtype}")
    check(lib().ocx_periodic_pad(_ptr(a), _ptr(out), lead, sy, sx, kid, hy, hx, elem, int(crop), _stream()),
          "ocx_periodic_pad")
    return out


def diffuse_h(T, Tw, h, act, cx, cy, okx, oky, V, kh, dt, out=None):
    """
    Explicit Laplacian diffusion along the layers, the ALE kernel of ``TracerDiffusion.diffuse_h``
    (``_diffuse_ale_kernel``, core/dynamics/sigma_baroclinic.py:533-561; core/components/tracers.py:420-446), on a
    stack: T, Tw (nt, nz, ny, nx) or (nz, ny, nx); h (nz, ny, nx) zeroed on inactive cells; act (ny, nx);
    cx = L_u/dx_u (ny, nx+1), cy = L_v/dy_v (ny+1, nx); okx, oky face masks; V (nz, ny, nx).  Bit-identical to the
    reference's arithmetic (``ocx_diffuse_h_f64``).
    """
    squeeze = T.dim() == 3
    T4, W4 = (T[None], Tw[None]) if squeeze else (T, Tw)
    nt, n, ny, nx = T4.shape
    _need(T4, (nt, n, ny, nx), "T")
    _need(W4, (nt, n, ny, nx), "Tw")
    _need(h, (n, ny, nx), "h")
    _need(V, (n, ny, nx), "V")
    _need(cx, (ny, nx + 1), "cx")
    _need(cy, (ny + 1, nx), "cy")
    a, ox, oy = _ok(act, (ny, nx)), _ok(okx, (ny, nx + 1)), _ok(oky, (ny + 1, nx))
    o4 = torch.empty_like(T4) if out is None else _need(out if not squeeze else out[None], T4.shape, "out")
    check(lib().ocx_diffuse_h_f64(_ptr(T4), _ptr(W4), _ptr(h), _ptr(a), _ptr(cx), _ptr(cy), _ptr(ox), _ptr(oy),
                                  _ptr(V), nt, n, ny, nx, float(kh), float(dt), _ptr(o4), _stream()),
          "ocx_diffuse_h_f64")
    return o4[0] if squeeze else o4


def diffuse_h_sigma(T, Tw, ds, gx, gy, okx, oky, V, kh, dt, out=None):
    """
    Explicit Laplacian diffusion along sigma surfaces, the sigma branch of ``TracerDiffusion.diffuse_h``
    (core/components/tracers.py:421-428, 450-481): F_x = -K_h G_u ds_k (Tw_i - Tw_i-1) with G_u = L_u D_u / dx_u
    (gx (ny, nx+1), D_u with NaN read as 0; gy likewise), masked by okx / oky; ds (n,) the layer fractions.  The
    kernel of ``diffuse_h`` with every face thickness ds_k: its arithmetic ((-K_h G) h) dTw is the sigma formula's
    term for term.  T, Tw (nt, n, ny, nx) or (n, ny, nx); V (n, ny, nx).
    """
    n, ny, nx = V.shape
    d = torch.as_tensor(ds, dtype=_F64, device=V.device).reshape(n, 1, 1)
    h = d.expand(n, ny, nx).contiguous()
    act = torch.ones((ny, nx), dtype=torch.uint8, device=V.device)
    return diffuse_h(T, Tw, h, act, gx, gy, okx, oky, V, kh, dt, out=out)


def vertical_implicit(u, dz, nu, r_bed, dt, tau_top=None, out=None, workspace: Workspace | None = None):
    """
    Backward-Euler vertical diffusion, column by column (``vertical_implicit`` / ``_thomas_columns``,
    core/dynamics/sigma_wd.py:57-99), same arithmetic: u (n, M) or a stack (nt, n, M); dz (n, M); nu (n+1, M);
    r_bed None, (M,) on layer 0 or (n, M); tau_top None or (M,).  ``out`` may be ``u`` (in place).
    """
    squeeze = u.dim() == 2
    u3 = u[None] if squeeze else u
    nt, n, M = u3.shape
    _need(u3, (nt, n, M), "u")
    _need(dz, (n, M), "dz")
    _need(nu, (n + 1, M), "nu")
    r_layers = 0
    if r_bed is not None:
        if r_bed.dim() == 1:
            _need(r_bed, (M,), "r_bed")
            r_layers = 1
        else:
            _need(r_bed, (n, M), "r_bed")
            r_layers = n
    if tau_top is not None:
        _need(tau_top, (M,), "tau_top")
    o3 = torch.empty_like(u3) if out is None else _need(out if not squeeze else out[None], u3.shape, "out")
    L = lib()
    ws = (workspace or _WS).get(int(L.ocx_vertical_implicit_workspace_bytes(n, M)))
    check(L.ocx_vertical_implicit_f64(_ptr(u3), _ptr(dz), _ptr(nu), _ptr(r_bed), r_layers, _ptr(tau_top), nt, n, M,
                                      float(dt), _ptr(o3), _ptr(ws), ws.numel(), _stream()), "ocx_vertical_implicit_f64")
    return o3[0] if squeeze else o3


def numerical_mixing(c, c2, V1, dt, out=None, accumulate=False):
    """Numerical mixing of the advection V1 (<c^2> - c^2) / dt per cell (``ocx_mix_num_f64``): c, c2 (nz, ny, nx)
    the advected c and c^2 (members of one stack); ``accumulate`` adds to ``out`` (the variance budget's +=).
    core/model/driver.py:905 (mix_num), core/components/variance_budget.py:60-62."""
    n, ny, nx = c.shape
    _need(c, (n, ny, nx), "c")
    _need(c2, (n, ny, nx), "c2")
    _need(V1, (n, ny, nx), "V1")
    if out is None:
        if accumulate:
            raise ValueError("accumulate needs out")
        out = torch.empty_like(c)
    _need(out, (n, ny, nx), "out")
    check(lib().ocx_mix_num_f64(_ptr(c), _ptr(c2), _ptr(V1), n, ny, nx, float(dt), int(bool(accumulate)), _ptr(out),
                                _stream()), "ocx_mix_num_f64")
    return out


def content(c, V):
    """Per-tracer content sum V c (FP64, deterministic) (``ocx_content_f64``)."""
    nt, n, ny, nx = c.shape
    _need(c, (nt, n, ny, nx), "c")
    _need(V, (n, ny, nx), "V")
    out = torch.empty((nt,), dtype=_F64, device=c.device)
    L = lib()
    ws = _WS_CONTENT.get(int(L.ocx_content_workspace_bytes(nt)))
    check(L.ocx_content_f64(_ptr(c), _ptr(V), nt, n, ny, nx, _ptr(out), _ptr(ws), ws.numel(), _stream()),
          "ocx_content_f64")
    return out
