"""
Drop-in ``TracerAdvection`` (role TRACER_ADVECTION) of pyOCEAN's 3-D core, running on the GPU.

Same constructor, attributes, method names, argument meanings, return types and errors as
``core/components/tracers.py:50-375``: NumPy arrays in (any layout or dtype, coerced to contiguous FP64 like the
reference), freshly allocated NumPy arrays out, no state kept from one call to the next except
``last_open_flux``.  Underneath, each call copies its inputs to HBM once, runs the HIP kernels of ``libocx.so``
and copies the result back; ``pyocean_amd.ops`` has the same operations on device tensors for callers that keep
their state resident.

Index order (tracer, k, j, i), k = 0 at the bed; faces Qx (N, ny, nx+1), Qy (N, ny+1, nx), Fz (N+1, ny, nx).

Scope: SSP-RK3 with the WENO5 schemes ('weno5' = WENO5-Z, 'weno5js'), with or without the FCT limiter, and the
'upwind3' / 'upwind1' faces of the unlimited step, open or periodic edges.  The face choice follows the reference's
dispatch: the RK3 step, its RHS and flux_x take UP3 / donor faces for a scheme that does not start with 'weno5'
(tracers.py:110-115, 139-164, 194-206); the FCT step and faces_h run WENO5-Z for 'weno5' and WENO5-JS for every other
scheme (tracers.py:166-176).  tracer_time 'pc' (advect_numpy, tracers.py:303-330) runs a centred half step and the
full step with those faces; rhs and flux_x take ``centred=True`` (_centred_faces, sigma_baroclinic.py:632-652).
"""
from __future__ import annotations

import numpy as np
import torch

from . import ops
from ._native import OPEN_E, OPEN_N, OPEN_S, OPEN_W
from .contract import Issue, Role
from .periodic import NONE, TRACER_HALO

_EDGE_BITS = {"west": OPEN_W, "east": OPEN_E, "south": OPEN_S, "north": OPEN_N}
SCHEMES = ("weno5", "weno5js", "upwind3", "upwind1")
TIMES = ("ssprk3", "pc")


class OpenEdges:
    """Minimal stand-in for the open-boundary component: only ``open`` (tuple of edge names) is read."""

    def __init__(self, open_=()):
        self.open = tuple(open_)


def _dev(a, dtype=np.float64):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to(ops.require_gpu())


def _host(t):
    return t.cpu().numpy()


def _empty_result(C, flux_face_i):
    """The step's result for a stack with no cells: C_new of C's shape, and the face column (ntr, N, ny)."""
    out = np.empty_like(C)
    if flux_face_i is None:
        return out
    return out, np.empty(C.shape[:3])


class TracerAdvection:
    name = "tracer_advection"
    role = Role.TRACER_ADVECTION
    reads = ("T", "tr", "Qx", "Qy", "Fz")
    writes = ("T", "tr")
    per = NONE                       # periodic axes (pyocean_amd.periodic), set by the model

    def __init__(self, *, scheme: str, time: str, dt: float, ob) -> None:
        """tracers.py:57-63."""
        self.tracer_scheme, self.tracer_time, self.dt, self.ob = scheme, time, dt, ob
        self.record_open_flux = False
        self.last_open_flux: np.ndarray | None = None
        self._ws = ops.Workspace()

    # ------------------------------------------------------------------ contract (tracers.py:104-108)
    def check(self, ctx) -> list[Issue]:
        return []

    def state_fields(self) -> dict[str, np.ndarray]:
        return {}

    # ------------------------------------------------------------------ helpers
    def _flags(self, unlimited=False, rk=False, centred=False) -> int:
        """Flag word of the open edges and the faces: ``centred`` the centred faces; ``rk`` (the RK3 step, its
        RHS, flux_x) gives a scheme that does not start with 'weno5' the UP1 ('upwind1') or UP3 faces
        (tracers.py:110-115); otherwise the WENO5 faces of faces_h, Z for 'weno5' and JS for every other scheme
        (tracers.py:166-176)."""
        m = 0
        for e in getattr(self.ob, "open", ()) or ():
            m |= _EDGE_BITS[e]
        if centred:
            return ops.flags_of(m, centred=True)
        sch = self.tracer_scheme
        if rk and not sch.startswith("weno5"):
            return ops.flags_of(m, up=1 if sch == "upwind1" else 3)
        return ops.flags_of(m, js=sch != "weno5", unlimited=unlimited)

    def _pad(self, t, kind):
        per = self.per
        return ops.periodic_pad(t, kind, TRACER_HALO, per.x, per.y)

    def _crop(self, t):
        per = self.per
        return ops.periodic_pad(t, "c", TRACER_HALO, per.x, per.y, crop=True)

    @staticmethod
    def _shapes(C, V0, V1, Qx, Qy, Fz=None):
        nt, n, ny, nx = C.shape
        for a, s, nm in ((V0, (n, ny, nx), "V0"), (V1, (n, ny, nx), "V1"), (Qx, (n, ny, nx + 1), "Qx"),
                         (Qy, (n, ny + 1, nx), "Qy")):
            if np.shape(a) != s:
                raise ValueError(f"{nm}: shape {np.shape(a)} vs tracer layers {C.shape[1:]} (expected {s})")
        if Fz is not None and np.shape(Fz) != (n + 1, ny, nx):
            raise ValueError(f"Fz: shape {np.shape(Fz)}, expected {(n + 1, ny, nx)}")

    # ------------------------------------------------------------------ the step with FCT (tracers.py:208-254)
    def advect_fct(self, C: np.ndarray, V0: np.ndarray, V1: np.ndarray, Qx: np.ndarray, Qy: np.ndarray,
                   ok: np.ndarray, unlimited: bool = False, inflow: np.ndarray | None = None,
                   flux_face_i: int | None = None):
        """
        SSP-RK3 WENO5 advection of the stack C (ntr, N, ny, nx) with Zalesak (1979) FCT, on the GPU.
        ``unlimited``: C = 1 (the RK3 result).  ``flux_face_i``: also return the step's effective flux through
        u-face column i, F_L + C A (F_H when unlimited), (ntr, N, ny): (C_new, fx_i).
        """
        C = np.ascontiguousarray(C, dtype=float)
        self._shapes(C, V0, V1, Qx, Qy)
        if C.size == 0:
            # an empty stack (no tracers, layers or cells): nothing to advect, as the reference's loops over none
            if self.record_open_flux and not unlimited:
                self.last_open_flux = np.zeros(C.shape[0])
            return _empty_result(C, flux_face_i)
        flags = self._flags(unlimited)
        okc = np.broadcast_to(ok, C.shape[1:])
        d = {k: _dev(v) for k, v in (("C", C), ("V0", V0), ("V1", V1), ("Qx", Qx), ("Qy", Qy))}
        d["ok"] = _dev(okc, np.uint8)
        d["inflow"] = None if inflow is None else _dev(inflow)
        fi = flux_face_i
        if self.per:
            H = TRACER_HALO
            d = {k: (None if v is None else self._pad(v, {"Qx": "u", "Qy": "v"}.get(k, "c"))) for k, v in d.items()}
            fi = None if fi is None else fi + (H if self.per.x else 0)
        want_open = self.record_open_flux and not unlimited
        res = ops.advect_fct(d["C"], d["V0"], d["V1"], d["Qx"], d["Qy"], d["ok"], float(self.dt), d["inflow"], flags,
                             flux_face_i=fi, open_flux=want_open, workspace=self._ws)
        out, fxi, of = (res, None, None) if not isinstance(res, tuple) else res
        if self.per:
            out = self._crop(out)
            if fxi is not None and self.per.y:
                fxi = fxi[..., TRACER_HALO:fxi.shape[-1] - TRACER_HALO].contiguous()
        if want_open:
            self.last_open_flux = _host(of)
        if flux_face_i is None:
            return _host(out)
        return _host(out), _host(fxi)

    # ------------------------------------------------------------------ the step without limiter (tracers.py:342-375)
    def advect(self, C: np.ndarray, V0: np.ndarray, V1: np.ndarray, Qx: np.ndarray, Qy: np.ndarray,
               Fz: np.ndarray, ok: np.ndarray, inflow: np.ndarray | None = None,
               flux_face_i: int | None = None):
        """
        SSP-RK3 flux-form advection of the stack C with frozen transports Qx, Qy and Fz; layer volumes go linearly
        from V0 to V1.  ``flux_face_i``: also return (F(c0) + F(c1))/6 + 2F(c2)/3 of u-face column i.
        tracer_time 'pc': the centred half step to c^{n+1/2}, then the full step with its faces; ``flux_face_i``
        then gives F(c^{n+1/2}) (tracers.py:324-330).
        """
        C = np.ascontiguousarray(C, dtype=float)
        if np.shape(V0) != C.shape[1:] or np.shape(V1) != C.shape[1:]:
            raise ValueError(f"_advect: volume shape {np.shape(V0)}/{np.shape(V1)} vs tracer layers {C.shape[1:]}")
        self._shapes(C, V0, V1, Qx, Qy, Fz)
        if C.size == 0:
            return _empty_result(C, flux_face_i)
        flags = self._flags(rk=True)
        vert = bool(np.any(np.asarray(Fz) != 0.0))
        okc = np.broadcast_to(ok, C.shape[1:])
        d = {k: _dev(v) for k, v in (("C", C), ("V0", V0), ("V1", V1), ("Qx", Qx), ("Qy", Qy), ("Fz", Fz))}
        d["ok"] = _dev(okc, np.uint8)
        d["inflow"] = None if inflow is None else _dev(inflow)
        fi = flux_face_i
        if self.per:
            H = TRACER_HALO
            d = {k: (None if v is None else self._pad(v, {"Qx": "u", "Qy": "v"}.get(k, "c"))) for k, v in d.items()}
            fi = None if fi is None else fi + (H if self.per.x else 0)
        step = ops.advect_pc if self.tracer_time == "pc" else ops.advect_rk3
        res = step(d["C"], d["V0"], d["V1"], d["Qx"], d["Qy"], d["Fz"], d["ok"], float(self.dt), d["inflow"], flags,
                   vert=vert, flux_face_i=fi, workspace=self._ws)
        out, fxi = (res, None) if fi is None else res
        if self.per:
            out = self._crop(out)
            if fxi is not None and self.per.y:
                fxi = fxi[..., TRACER_HALO:fxi.shape[-1] - TRACER_HALO].contiguous()
        if flux_face_i is None:
            return _host(out)
        return _host(out), _host(fxi)

    # ------------------------------------------------------------------ single-stage helpers (tracers.py:153-206)
    def faces_h(self, c: np.ndarray, Qx: np.ndarray, Qy: np.ndarray, ok: np.ndarray, ext: np.ndarray | None = None,
                inflow: np.ndarray | None = None):
        """WENO5 horizontal face fluxes (fx, fy) of the stack; ``ext`` the exterior values (default inflow, else c)."""
        c = np.ascontiguousarray(c, dtype=float)
        nt, n, ny, nx = c.shape
        e = (None if inflow is None else inflow) if ext is None else ext
        fx, fy = ops.faces_h(_dev(c), _dev(Qx), _dev(Qy), _dev(np.broadcast_to(ok, c.shape[1:]), np.uint8),
                             None if e is None else _dev(e), self._flags())
        return _host(fx), _host(fy)

    def faces(self, c: np.ndarray, flux: np.ndarray, ok: np.ndarray, axis: int, centred: bool = False) -> np.ndarray:
        """Face values of c along ``axis`` (tracers.py:110-115): the centred faces (``centred``), the WENO5 faces
        ('weno5' Z, 'weno5js' JS) or the UP3 / donor faces ('upwind1') of every other scheme, upwind-ordered by the
        sign of ``flux``."""
        if centred:
            return centred_faces(c, ok, axis, n_out=np.shape(flux)[axis])
        if self.tracer_scheme.startswith("weno5"):
            return weno5_faces(c, flux, ok, axis, variant="js" if self.tracer_scheme == "weno5js" else "z")
        return up3_faces(c, flux, ok, axis, order=1 if self.tracer_scheme == "upwind1" else 3)

    def open_faces(self, face: np.ndarray, c: np.ndarray, Q: np.ndarray, axis: int,
                   inflow: np.ndarray | None = None) -> np.ndarray:
        """
        Tracer values on the open-boundary faces (tracers.py:117-137): upstream - the boundary cell where the flow
        leaves, the exterior value (``inflow``, default c) where it enters - on the open edges of the axis (3: west
        and east, 2: south and north) of a face stack ``face`` (ntr, N, ny, nx+1) or (ntr, N, ny+1, nx); a new
        array, the other faces copied.  Without open edges ``face`` itself is returned.
        """
        if not self.ob.open:
            return face
        m = 0
        for e in self.ob.open:
            m |= _EDGE_BITS[e]
        c = np.ascontiguousarray(c, dtype=float)
        ax = 0 if axis == 3 else 1
        f = _dev(face)
        ops.open_faces(f, _dev(c), _dev(Q), ax, m, None if inflow is None else _dev(inflow))
        return _host(f)

    def upwind_faces_h(self, c: np.ndarray, Qx: np.ndarray, Qy: np.ndarray, inflow: np.ndarray | None = None):
        """Donor-cell face fluxes Q c_up (the FCT low-order flux); open edges take the inflow values on inflow."""
        c = np.ascontiguousarray(c, dtype=float)
        fx, fy = ops.upwind_faces_h(_dev(c), _dev(Qx), _dev(Qy), None if inflow is None else _dev(inflow),
                                    self._flags() & 15)
        return _host(fx), _host(fy)

    def rhs(self, c: np.ndarray, Qx: np.ndarray, Qy: np.ndarray, Fz: np.ndarray, ok: np.ndarray,
            centred: bool = False, inflow: np.ndarray | None = None) -> np.ndarray:
        """Advection RHS -div(Q c): WENO faces on x, y and (any Fz != 0) z, or the UP3 / UP1 / centred
        (``centred``) faces of rhs_numpy on x, y and z (tracers.py:139-164)."""
        c = np.ascontiguousarray(c, dtype=float)
        vert = bool(np.any(np.asarray(Fz) != 0.0))
        flags = self._flags(centred=True) if centred else self._flags(rk=True)
        r = ops.rhs(_dev(c), _dev(Qx), _dev(Qy), _dev(Fz), _dev(np.broadcast_to(ok, c.shape[1:]), np.uint8),
                    None if inflow is None else _dev(inflow), flags, vert=vert)
        return _host(r)

    def flux_x(self, c: np.ndarray, Qx: np.ndarray, Qy: np.ndarray, ok: np.ndarray, centred: bool = False,
               inflow: np.ndarray | None = None) -> np.ndarray:
        """x-face fluxes Q c_face (ntr, N, ny, nx+1) that ``rhs`` differences (the TEF transect flux)."""
        if not centred and self.tracer_scheme.startswith("weno5"):
            return self.faces_h(c, Qx, Qy, np.broadcast_to(ok, np.shape(c)[1:]), inflow=inflow)[0]
        c = np.ascontiguousarray(c, dtype=float)
        flags = self._flags(centred=True) if centred else self._flags(rk=True)
        fx, _ = ops.faces_h(_dev(c), _dev(Qx), _dev(Qy), _dev(np.broadcast_to(ok, c.shape[1:]), np.uint8),
                            None if inflow is None else _dev(inflow), flags)
        return _host(fx)

    def content(self, C: np.ndarray, V: np.ndarray) -> np.ndarray:
        """Per-tracer content sum V c in FP64 (deterministic order)."""
        C = np.ascontiguousarray(C, dtype=float)
        return _host(ops.content(_dev(C), _dev(V)))

    # ------------------------------------------------------------------ the reference's NumPy twins (tracers.py:139,
    # 256, 303): same signatures; here they run the same GPU kernels as advect_fct, advect and rhs.  The reference's
    # tests hold its twins to its fused kernels within 1e-13 (FCT) and 1e-12 (RHS) and bit for bit (RK3); the GPU
    # kernels equal the fused ones bit for bit.
    def advect_fct_numpy(self, C, V0, V1, Qx, Qy, ok, unlimited=False, inflow=None, flux_face_i=None):
        """``advect_fct`` (tracers.py:256-301 is the reference's NumPy twin)."""
        return self.advect_fct(C, V0, V1, Qx, Qy, ok, unlimited=unlimited, inflow=inflow, flux_face_i=flux_face_i)

    def advect_numpy(self, C, V0, V1, Qx, Qy, Fz, ok, inflow=None, flux_face_i=None):
        """``advect`` (tracers.py:303-340 is the reference's NumPy twin, and its 'pc' path)."""
        return self.advect(C, V0, V1, Qx, Qy, Fz, ok, inflow=inflow, flux_face_i=flux_face_i)

    def rhs_numpy(self, c, Qx, Qy, Fz, ok, centred=False, inflow=None):
        """``rhs`` (tracers.py:139-151 is the reference's vectorised twin)."""
        return self.rhs(c, Qx, Qy, Fz, ok, centred=centred, inflow=inflow)


# ---------------------------------------------------------------------------- face values along any axis (module level)
def _faces_call(c, flux, ok, axis, flags):
    c = np.asarray(c, dtype=float)
    flux = np.asarray(flux, dtype=float)
    okb = np.asarray(ok, dtype=bool)
    return _host(ops.faces(_dev(c), _dev(flux), _dev(okb, np.uint8), axis, flags))


def weno5_faces(c: np.ndarray, flux: np.ndarray, ok: np.ndarray, axis: int, variant: str = "z") -> np.ndarray:
    """WENO5 face values of c along ``axis`` on the GPU, as ``_weno5_faces`` (sigma_baroclinic.py:564-578): faces
    1 .. n-1 upwind-ordered by the sign of ``flux`` (which sets the face count), the others 0; variant 'z' (WENO5-Z)
    or 'js' (WENO5-JS), anything else a ValueError."""
    if variant not in ("z", "js"):
        raise ValueError(f"unknown WENO variant {variant!r}")
    return _faces_call(c, flux, ok, axis, ops.flags_of(js=variant == "js"))


def up3_faces(c: np.ndarray, flux: np.ndarray, ok: np.ndarray, axis: int, order: int = 3) -> np.ndarray:
    """UP3 (or, ``order`` 1, donor-cell) face values along ``axis`` on the GPU, as ``_up3_faces``
    (sigma_baroclinic.py:164-187)."""
    return _faces_call(c, flux, ok, axis, ops.flags_of(up=1 if order == 1 else 3))


def centred_faces(c: np.ndarray, ok: np.ndarray, axis: int, n_out: int | None = None) -> np.ndarray:
    """Centred 6th/4th/2nd-order face values along ``axis`` on the GPU, as ``_centred_faces``
    (sigma_baroclinic.py:632-652); ``n_out`` faces (default n+1)."""
    c = np.asarray(c, dtype=float)
    n = c.shape[axis]
    shape = list(c.shape)
    shape[axis] = n + 1 if n_out is None else n_out
    return _faces_call(c, np.zeros(shape), ok, axis, ops.flags_of(centred=True))
