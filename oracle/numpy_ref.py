"""
Vectorised NumPy restatement of pyOCEAN's tracer advection (test infrastructure only; see ``oracle/__init__``).

Index order (tracer t, layer k, row j, column i), k = 0 at the bed.  Face arrays: Qx (N, ny, nx+1),
Qy (N, ny+1, nx), Fz (N+1, ny, nx); cells V0, V1, ok (N, ny, nx); stacks (nt, N, ny, nx).  All FP64.

Every expression keeps the operand order of the reference so that IEEE rounding matches it; the reference
file:line each function follows is given in its docstring.
"""
from __future__ import annotations

import numpy as np

C1312 = 13 / 12          # the constant 13/12 of the smoothness indicators, as Python folds it
EPS_Z = 1e-40            # WENO-Z epsilon (sigma_baroclinic.py:219)
EPS_JS = 1e-6            # WENO-JS and WENO3 epsilon (sigma_baroclinic.py:224, 229-230)

OPEN_W, OPEN_E, OPEN_S, OPEN_N = 1, 2, 4, 8


def edges_mask(open_edges) -> int:
    """Bit mask of the open edge names ('west', 'east', 'south', 'north')."""
    bits = {"west": OPEN_W, "east": OPEN_E, "south": OPEN_S, "north": OPEN_N}
    m = 0
    for e in open_edges:
        m |= bits[e]
    return m


# ------------------------------------------------------------------------------------------------ WENO values
def weno_value(v1, v2, v3, v4, v5, m5, m3, z_variant=True):
    """
    Face value from the upwind-ordered stencil v1..v5 (v3 the upwind cell): WENO5-Z or -JS where ``m5`` (all five
    cells valid), WENO3 where ``m3`` (v2..v4 valid), else v3.  Follows ``_wval`` (sigma_baroclinic.py:204-233)
    and ``_weno5_faces_numpy.recon`` (:603-618), term for term.
    """
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        q0 = (2 * v1 - 7 * v2 + 11 * v3) / 6
        q1 = (-v2 + 5 * v3 + 2 * v4) / 6
        q2 = (2 * v3 + 5 * v4 - v5) / 6
        b0 = C1312 * (v1 - 2 * v2 + v3) ** 2 + 0.25 * (v1 - 4 * v2 + 3 * v3) ** 2
        b1 = C1312 * (v2 - 2 * v3 + v4) ** 2 + 0.25 * (v2 - v4) ** 2
        b2 = C1312 * (v3 - 2 * v4 + v5) ** 2 + 0.25 * (3 * v3 - 4 * v4 + v5) ** 2
        if z_variant:
            tau = np.abs(b2 - b0)
            a0 = 0.1 * (1 + tau / (EPS_Z + b0))
            a1 = 0.6 * (1 + tau / (EPS_Z + b1))
            a2 = 0.3 * (1 + tau / (EPS_Z + b2))
        else:
            a0 = 0.1 / (EPS_JS + b0) ** 2
            a1 = 0.6 / (EPS_JS + b1) ** 2
            a2 = 0.3 / (EPS_JS + b2) ** 2
        w5 = (a0 * q0 + a1 * q1 + a2 * q2) / (a0 + a1 + a2)
        c0 = (-v2 + 3 * v3) / 2
        c1 = (v3 + v4) / 2
        e0 = (1 / 3) / (EPS_JS + (v3 - v2) ** 2) ** 2
        e1 = (2 / 3) / (EPS_JS + (v4 - v3) ** 2) ** 2
        w3 = (e0 * c0 + e1 * c1) / (e0 + e1)
    return np.where(m5, w5, np.where(m3, w3, v3))


def _shifted(c, ok, s):
    """c[..., m + s] and its validity along the last axis; positions outside the row are 0 and invalid."""
    n = c.shape[-1]
    val = np.zeros_like(c)
    good = np.zeros(ok.shape, bool)
    lo, hi = max(0, -s), min(n, n - s)
    if hi > lo:
        val[..., lo:hi] = c[..., lo + s:hi + s]
        good[..., lo:hi] = ok[..., lo + s:hi + s]
    return val, good


def weno5_faces(c, flux, ok, axis, z_variant=True):
    """
    WENO face values on the interior faces 1..n-1 along ``axis`` (face m between cells m-1 and m), upwind with
    respect to the sign of ``flux`` (>= 0: from the low side); other faces 0.  ``flux`` has the face count of the
    output along ``axis`` (n+1 for x/y, n for the vertical as ``rhs_numpy`` passes it).  Follows
    ``_weno5_faces_numpy`` (sigma_baroclinic.py:581-629) and ``_weno5_rows`` (:236-262).
    """
    cm = np.moveaxis(np.asarray(c, float), axis, -1)
    om = np.moveaxis(np.asarray(ok, bool), axis, -1)
    fm = np.moveaxis(np.asarray(flux, float), axis, -1)
    lead = np.broadcast_shapes(cm.shape[:-1], om.shape[:-1], fm.shape[:-1])
    n = cm.shape[-1]
    cm = np.broadcast_to(cm, lead + (n,))
    om = np.broadcast_to(om, lead + (n,))
    fm = np.broadcast_to(fm, lead + (fm.shape[-1],))
    sh = {s: _shifted(cm, om, s) for s in (-3, -2, -1, 0, 1, 2)}
    v = {s: sh[s][0] for s in sh}
    k = {s: sh[s][1] for s in sh}
    pos = weno_value(v[-3], v[-2], v[-1], v[0], v[1], k[-3] & k[-2] & k[-1] & k[0] & k[1],
                     k[-2] & k[-1] & k[0], z_variant)
    neg = weno_value(v[2], v[1], v[0], v[-1], v[-2], k[2] & k[1] & k[0] & k[-1] & k[-2],
                     k[1] & k[0] & k[-1], z_variant)
    face = np.zeros(fm.shape)
    face[..., 1:n] = np.where(fm[..., 1:n] >= 0, pos[..., 1:n], neg[..., 1:n])
    return np.moveaxis(face, -1, axis)


# ------------------------------------------------------------------------------------------------ face fluxes
def faces_h(c, Qx, Qy, ok, ext, open_mask=0, z_variant=True):
    """
    Horizontal face fluxes (fx, fy) = Q * c_face: WENO faces on interior faces with Q != 0 (Q == 0: exactly 0),
    upstream values on open edges (``ext`` on inflow, the edge cell on outflow), 0 on closed edges.  Follows
    ``_advect_faces_h`` (sigma_baroclinic.py:265-321).
    """
    c = np.asarray(c, float)
    nt, n, ny, nx = c.shape
    okb = np.broadcast_to(ok, c.shape)
    Qxb = np.broadcast_to(Qx, (nt,) + Qx.shape)
    Qyb = np.broadcast_to(Qy, (nt,) + Qy.shape)
    fx = np.where(Qxb == 0.0, 0.0, Qxb * weno5_faces(c, Qxb, okb, 3, z_variant))
    fy = np.where(Qyb == 0.0, 0.0, Qyb * weno5_faces(c, Qyb, okb, 2, z_variant))
    fx[..., 0] = 0.0
    fx[..., nx] = 0.0
    fy[..., 0, :] = 0.0
    fy[..., ny, :] = 0.0
    _open_edges(fx, fy, c, Qx, Qy, ext, open_mask, donor=False)
    return fx, fy


def _open_edges(fx, fy, c, Qx, Qy, ext, open_mask, donor):
    """Open-edge face fluxes q * (ext on inflow, else the edge cell) (sigma_baroclinic.py:298-303, 315-320)."""
    if open_mask & OPEN_W:
        q = Qx[None, ..., 0]
        fx[..., 0] = q * np.where(q > 0.0, ext[..., 0], c[..., 0])
    if open_mask & OPEN_E:
        q = Qx[None, ..., -1]
        fx[..., -1] = q * np.where(q < 0.0, ext[..., -1], c[..., -1])
    if open_mask & OPEN_S:
        q = Qy[None, :, 0, :]
        fy[..., 0, :] = q * np.where(q > 0.0, ext[..., 0, :], c[..., 0, :])
    if open_mask & OPEN_N:
        q = Qy[None, :, -1, :]
        fy[..., -1, :] = q * np.where(q < 0.0, ext[..., -1, :], c[..., -1, :])


def open_faces(face, c, Q, axis, open_mask, ext=None):
    """
    Open-edge face VALUES (not fluxes) of one axis of a face stack: ``axis`` 3 the x faces (west, east), 2 the y
    faces (south, north); on each open edge the exterior value (``ext``, default c) of the boundary cell where the
    flux enters (Q > 0 on the low edge, Q < 0 on the high edge), the boundary cell otherwise; a copy, other faces
    unchanged.  Follows ``TracerAdvection.open_faces`` (tracers.py:117-137).
    """
    ext = c if ext is None else ext
    face = np.array(face, dtype=float, copy=True)
    lo, hi = (OPEN_W, OPEN_E) if axis == 3 else (OPEN_S, OPEN_N)
    for bit, f_i, c_i, sign in ((lo, 0, 0, 1.0), (hi, -1, -1, -1.0)):
        if not open_mask & bit:
            continue
        q = np.take(Q, f_i, axis=axis - 1)
        cell = np.take(c, c_i, axis=axis)
        out = np.take(ext, c_i, axis=axis)
        val = np.where(sign * q > 0, out, cell)
        if axis == 3:
            face[..., f_i] = val
        else:
            face[..., f_i, :] = val
    return face


def upwind_faces_h(c, Qx, Qy, ext, open_mask=0):
    """
    Donor-cell face fluxes F_L = Q * c_up on all interior faces (Q >= 0: the low-side cell), open edges as
    ``faces_h``, closed edges 0.  Follows ``TracerAdvection.upwind_faces_h`` (tracers.py:178-193) and
    ``_fct_fluxes`` (sigma_baroclinic.py:482-505).
    """
    c = np.asarray(c, float)
    nt, n, ny, nx = c.shape
    fx = np.zeros((nt, n, ny, nx + 1))
    fy = np.zeros((nt, n, ny + 1, nx))
    qx = Qx[None, ..., 1:-1]
    fx[..., 1:-1] = qx * np.where(qx >= 0, c[..., :-1], c[..., 1:])
    qy = Qy[None, :, 1:-1, :]
    fy[..., 1:-1, :] = qy * np.where(qy >= 0, c[..., :-1, :], c[..., 1:, :])
    _open_edges(fx, fy, c, Qx, Qy, ext, open_mask, donor=True)
    return fx, fy


def faces_z(c, Fz, ok, z_variant=True):
    """
    Vertical face fluxes fz (nt, N+1, ny, nx) = Fz * c_face on the interfaces 1..N-1 (Fz == 0 skipped, bed and
    surface 0), WENO along the layer index with each layer's own ``ok``.  Follows ``_advect_div``'s vertical pass
    (sigma_baroclinic.py:335-353).
    """
    c = np.asarray(c, float)
    nt, n, ny, nx = c.shape
    okb = np.broadcast_to(ok, c.shape)
    Fzb = np.broadcast_to(Fz[:-1], c.shape)
    v = weno5_faces(c, Fzb, okb, 1, z_variant)
    fz = np.zeros((nt, n + 1, ny, nx))
    fz[:, :-1] = np.where(Fzb == 0.0, 0.0, Fzb * v)
    fz[:, 0] = 0.0
    return fz


def div_h(fx, fy):
    """((fx_e - fx_w) + fy_n) - fy_s, the grouping of _fct_stage (sigma_baroclinic.py:463) and the NumPy oracle."""
    return fx[..., 1:] - fx[..., :-1] + fy[..., 1:, :] - fy[..., :-1, :]


def rhs(c, Qx, Qy, Fz, ok, ext, open_mask=0, z_variant=True, vert=None):
    """
    Advection RHS -div(Q c) of the stack, x, y and (if ``vert``, default: any Fz != 0) z faces.  Follows
    ``_advect_div`` (sigma_baroclinic.py:324-364) and ``TracerAdvection.rhs`` (tracers.py:153-164).
    """
    fx, fy = faces_h(c, Qx, Qy, ok, ext, open_mask, z_variant)
    if vert is None:
        vert = bool(np.any(np.asarray(Fz) != 0.0))
    if vert:
        fz = faces_z(c, Fz, ok, z_variant)
        return -(fx[..., 1:] - fx[..., :-1] + fy[..., 1:, :] - fy[..., :-1, :] + fz[:, 1:] - fz[:, :-1])
    return -(fx[..., 1:] - fx[..., :-1] + fy[..., 1:, :] - fy[..., :-1, :])


def up_faces(c, flux, ok, axis, order=3):
    """
    UP3 (``order`` 3) or donor-cell (``order`` 1) face values - not yet times the flux - on the faces along
    ``axis``: face f (1 <= f <= n-1, between cells f-1 and f) takes, for flux >= 0, (-c[f-2] + 5 c[f-1] + 2 c[f]) / 6
    where cells f-2, f-1, f are all ok, else c[f-1]; for flux < 0 (or NaN), (2 c[f-1] + 5 c[f] - c[f+1]) / 6 where
    cells f-1, f, f+1 are ok, else c[f]; cells outside 0..n-1 count as not ok.  Faces 0 and the last are 0.
    ``flux`` has n+1 faces (x, y) or n (z: Fz without the surface).  Follows ``_up3_faces``
    (sigma_baroclinic.py:164-187).
    """
    cm = np.moveaxis(np.asarray(c, float), axis, -1)
    om = np.moveaxis(np.asarray(ok).astype(bool), axis, -1)
    qm = np.moveaxis(np.asarray(flux, float), axis, -1)
    n = cm.shape[-1]
    out = np.zeros(qm.shape)
    if n >= 2:
        lo, hi = cm[..., :-1], cm[..., 1:]
        qf = qm[..., 1:n]
        if order == 1:
            val = np.where(qf >= 0, lo, hi)
        else:
            pad_c = np.zeros(cm.shape[:-1] + (1,))
            pad_o = np.zeros(om.shape[:-1] + (1,), bool)
            ce = np.concatenate([pad_c, cm, pad_c], axis=-1)            # ce[..., m + 1] = cell m
            oe = np.concatenate([pad_o, om, pad_o], axis=-1)
            f = np.arange(1, n)
            olo, ohi = om[..., :-1], om[..., 1:]
            with np.errstate(invalid="ignore", over="ignore"):
                pos = np.where(oe[..., f - 1] & olo & ohi, (-ce[..., f - 1] + 5.0 * lo + 2.0 * hi) / 6.0, lo)
                neg = np.where(olo & ohi & oe[..., f + 2], (2.0 * lo + 5.0 * hi - ce[..., f + 2]) / 6.0, hi)
            val = np.where(qf >= 0, pos, neg)
        out[..., 1:n] = val
    return np.moveaxis(out, -1, axis)


def centred_faces(c, ok, axis, n_out=None):
    """
    Centred face values on the faces along ``axis``, whatever the flux: face f (1 <= f <= n-1) takes
    (37 (c[f-1] + c[f]) - 8 (c[f-2] + c[f+1]) + (c[f-3] + c[f+2])) / 60 where cells f-3 .. f+2 are all ok, else
    (7 (c[f-1] + c[f]) - (c[f-2] + c[f+1])) / 12 where cells f-2 .. f+1 are ok, else (c[f-1] + c[f]) / 2; cells
    outside 0..n-1 count as not ok; faces 0 and the last are 0.  ``n_out``: number of faces (default n+1).
    Follows ``_centred_faces`` (sigma_baroclinic.py:632-652).
    """
    cm = np.moveaxis(np.asarray(c, float), axis, -1)
    om = np.moveaxis(np.asarray(ok).astype(bool), axis, -1)
    n = cm.shape[-1]
    out = np.zeros(cm.shape[:-1] + (n + 1 if n_out is None else n_out,))
    if n >= 2:
        ce = np.concatenate([np.zeros(cm.shape[:-1] + (3,)), cm, np.zeros(cm.shape[:-1] + (3,))], axis=-1)
        oe = np.concatenate([np.zeros(om.shape[:-1] + (3,), bool), om, np.zeros(om.shape[:-1] + (3,), bool)],
                            axis=-1)
        f = np.arange(1, n) + 3                                          # ce[..., m + 3] = cell m
        a = ce[..., f - 1] + ce[..., f]
        b = ce[..., f - 2] + ce[..., f + 1]
        m4 = oe[..., f - 2] & oe[..., f - 1] & oe[..., f] & oe[..., f + 1]
        m6 = m4 & oe[..., f - 3] & oe[..., f + 2]
        f6 = (37.0 * a - 8.0 * b + (ce[..., f - 3] + ce[..., f + 2])) / 60.0
        f4 = (7.0 * a - b) / 12.0
        out[..., 1:n] = np.where(m6, f6, np.where(m4, f4, 0.5 * a))
    return np.moveaxis(out, -1, axis)


def _np_faces(c, flux, ok, axis, order):
    """Face values of rhs_numpy's ``faces`` (tracers.py:110-115): ``order`` 'c' centred, else UP3 / UP1."""
    if order == "c":
        return centred_faces(c, ok, axis, n_out=flux.shape[axis])
    return up_faces(c, flux, ok, axis, order)


def faces_up_h(c, Qx, Qy, ok, ext, open_mask=0, order=3):
    """
    Horizontal face fluxes (fx, fy) = Q * face of the UP3 / UP1 (``order`` 3 / 1) or centred (``order`` 'c')
    faces: every face, interior and closed edge, is Q times its face value (a closed edge Q * 0); open edges as
    ``faces_h``.  Follows ``rhs_numpy`` and ``open_faces`` (tracers.py:117-151) and ``flux_x`` (tracers.py:194-206).
    """
    c = np.asarray(c, float)
    nt = c.shape[0]
    okb = np.broadcast_to(ok, c.shape)
    Qxb = np.broadcast_to(Qx, (nt,) + Qx.shape)
    Qyb = np.broadcast_to(Qy, (nt,) + Qy.shape)
    fx = Qxb * _np_faces(c, Qxb, okb, 3, order)
    fy = Qyb * _np_faces(c, Qyb, okb, 2, order)
    _open_edges(fx, fy, c, Qx, Qy, ext, open_mask, donor=False)
    return fx, fy


def rhs_up(c, Qx, Qy, Fz, ok, ext, open_mask=0, order=3):
    """
    Advection RHS -div(Q c) of the stack with UP3 / UP1 / centred (``order`` 3, 1, 'c') faces on x, y and z, the
    z faces always formed: fz[k] = Fz[k] * face_k for k < N (face_0 = 0), fz[N] = 0.  Follows ``rhs_numpy``
    (tracers.py:139-151).
    """
    c = np.asarray(c, float)
    nt, n, ny, nx = c.shape
    fx, fy = faces_up_h(c, Qx, Qy, ok, ext, open_mask, order)
    Fzb = np.broadcast_to(np.asarray(Fz, float)[:-1], c.shape)
    fz = np.zeros((nt, n + 1, ny, nx))
    fz[:, :-1] = Fzb * _np_faces(c, Fzb, np.broadcast_to(ok, c.shape), 1, order)
    return -(fx[..., 1:] - fx[..., :-1] + fy[..., 1:, :] - fy[..., :-1, :] + fz[:, 1:] - fz[:, :-1])


# ------------------------------------------------------------------------------------------------ FCT
def fct_coefficients(c0, cL, V1, Ax, Ay, ok, dt):
    """
    Zalesak (1979, eqs. 7-13) limiting coefficients (Cx, Cy) in [0, 1] of the antidiffusive fluxes A, with the
    bounds taken over the cell, its x and y neighbours and the layers k +- 1 (``ok`` cells only); cells not ok get
    R+- = 0; edge faces 0.  Follows ``fct_coefficients_numpy`` (sigma_baroclinic.py:367-404) and
    ``_fct_coeff_kernel`` (:407-449).
    """
    okb = np.broadcast_to(ok, c0.shape)
    hi = np.where(okb, np.maximum(c0, cL), -np.inf)
    lo = np.where(okb, np.minimum(c0, cL), np.inf)

    def around(a, f):
        out = a.copy()
        out[..., 1:] = f(out[..., 1:], a[..., :-1])
        out[..., :-1] = f(out[..., :-1], a[..., 1:])
        out[..., 1:, :] = f(out[..., 1:, :], a[..., :-1, :])
        out[..., :-1, :] = f(out[..., :-1, :], a[..., 1:, :])
        out[:, 1:] = f(out[:, 1:], a[:, :-1])
        out[:, :-1] = f(out[:, :-1], a[:, 1:])
        return out

    cmax, cmin = around(hi, np.maximum), around(lo, np.minimum)
    p_in = dt * (np.maximum(Ax[..., :-1], 0.0) - np.minimum(Ax[..., 1:], 0.0)
                 + np.maximum(Ay[..., :-1, :], 0.0) - np.minimum(Ay[..., 1:, :], 0.0))
    p_out = dt * (np.maximum(Ax[..., 1:], 0.0) - np.minimum(Ax[..., :-1], 0.0)
                  + np.maximum(Ay[..., 1:, :], 0.0) - np.minimum(Ay[..., :-1, :], 0.0))
    with np.errstate(invalid="ignore"):
        q_in = np.maximum(V1[None] * (cmax - cL), 0.0)
        q_out = np.maximum(V1[None] * (cL - cmin), 0.0)
    r_in = np.where(okb & (p_in > 0), np.minimum(1.0, q_in / np.where(p_in > 0, p_in, 1.0)), 0.0)
    r_out = np.where(okb & (p_out > 0), np.minimum(1.0, q_out / np.where(p_out > 0, p_out, 1.0)), 0.0)
    Cx = np.zeros_like(Ax)
    Cy = np.zeros_like(Ay)
    Cx[..., 1:-1] = np.where(Ax[..., 1:-1] >= 0, np.minimum(r_in[..., 1:], r_out[..., :-1]),
                             np.minimum(r_in[..., :-1], r_out[..., 1:]))
    Cy[..., 1:-1, :] = np.where(Ay[..., 1:-1, :] >= 0, np.minimum(r_in[..., 1:, :], r_out[..., :-1, :]),
                                np.minimum(r_in[..., :-1, :], r_out[..., 1:, :]))
    return Cx, Cy


def _safe(q, v, C):
    """c = q / V where V > 0, else the step's initial C (tracers.py:275)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(v > 0, q / np.where(v > 0, v, 1.0), C)


def advect_fct(C, V0, V1, Qx, Qy, ok, dt, inflow=None, open_mask=0, z_variant=True, unlimited=False,
               flux_face_i=None, open_flux=False):
    """
    One SSP-RK3 + Zalesak FCT step of the stack C.  Follows ``advect_fct`` / ``advect_fct_numpy``
    (tracers.py:208-301): stage faces F0, F1, F2 with the stage's own values as exterior when ``inflow`` is None
    (every stage uses ``inflow`` when given), F_H = (F0 + F1)/6 + 2 F2/3, donor-cell F_L, A = F_H - F_L, c_L, the
    coefficients and V1 c = V0 c0 - dt div(F_L + C A).  ``unlimited``: C = 1.  Returns ``out`` or a dict with
    ``out`` and the optional side outputs ``fx_i`` (effective flux of u-face column i, (nt, N, ny)) and
    ``open_flux`` (net open-edge inflow rate per tracer, (nt,)).
    """
    C = np.ascontiguousarray(C, dtype=float)
    Qx, Qy = np.asarray(Qx, float), np.asarray(Qy, float)
    V0, V1 = np.asarray(V0, float), np.asarray(V1, float)
    okc = np.broadcast_to(ok, C.shape[1:])
    ext0 = C if inflow is None else np.asarray(inflow, float)
    dV = V1 - V0
    Va, Vb = V0 + dV, V0 + 0.5 * dV
    q0 = V0 * C
    fx0, fy0 = faces_h(C, Qx, Qy, okc, ext0, open_mask, z_variant)
    q1 = 0.0 * q0 + 1.0 * (q0 - dt * div_h(fx0, fy0))
    c1 = _safe(q1, Va, C)
    fx1, fy1 = faces_h(c1, Qx, Qy, okc, c1 if inflow is None else ext0, open_mask, z_variant)
    q2 = 0.75 * q0 + 0.25 * (q1 - dt * div_h(fx1, fy1))
    c2 = _safe(q2, Vb, C)
    fx2, fy2 = faces_h(c2, Qx, Qy, okc, c2 if inflow is None else ext0, open_mask, z_variant)
    FHx = (fx0 + fx1) / 6.0 + 2.0 * fx2 / 3.0
    FHy = (fy0 + fy1) / 6.0 + 2.0 * fy2 / 3.0
    FLx, FLy = upwind_faces_h(C, Qx, Qy, ext0, open_mask)
    Ax, Ay = FHx - FLx, FHy - FLy
    if unlimited:
        Gx, Gy = FHx, FHy
        out = _safe(q0 - dt * div_h(FHx, FHy), V1, C)
    else:
        cL = _safe(0.0 * q0 + 1.0 * (q0 - dt * div_h(FLx, FLy)), V1, C)
        Cx, Cy = fct_coefficients(C, cL, V1, Ax, Ay, okc, dt)
        Gx, Gy = FLx + Cx * Ax, FLy + Cy * Ay
        out = _safe(q0 - dt * div_h(Gx, Gy), V1, C)
    if flux_face_i is None and not open_flux:
        return out
    res = {"out": out}
    if flux_face_i is not None:
        res["fx_i"] = np.ascontiguousarray(Gx[..., flux_face_i])
    if open_flux:
        res["open_flux"] = open_inflow(Gx, Gy, open_mask)
    return res


def open_inflow(Fx, Fy, open_mask):
    """Net inflow rate per stack member, sum_W F - sum_E F + sum_S F - sum_N F (tracers.py:65-78)."""
    out = np.zeros(Fx.shape[0])
    for bit, F, sl, sgn in ((OPEN_W, Fx, np.s_[..., 0], 1.0), (OPEN_E, Fx, np.s_[..., -1], -1.0),
                            (OPEN_S, Fy, np.s_[..., 0, :], 1.0), (OPEN_N, Fy, np.s_[..., -1, :], -1.0)):
        if open_mask & bit:
            out += sgn * F[sl].reshape(F.shape[0], -1).sum(axis=1)
    return out


def advect_rk3(C, V0, V1, Qx, Qy, Fz, ok, dt, inflow=None, open_mask=0, z_variant=True, flux_face_i=None,
               up_order=0):
    """
    One SSP-RK3 step without limiter, x, y and z sweeps.  Follows ``advect`` (tracers.py:342-375) with
    ``_rk_stage`` (sigma_baroclinic.py:856-875): q1 = q0 + dt r0, q2 = 0.75 q0 + 0.25 (q1 + dt r1),
    q3 = q0/3 + 2/3 (q2 + dt r2), each with the ``safe`` division by Va, Vb, V1.  The exterior values are
    ``inflow`` in every stage, else each stage's own values.  ``up_order`` 3 or 1: the UP3 / UP1 faces of
    ``rhs_up`` instead of WENO.  Returns ``out`` or (out, fx_i).
    """
    C = np.ascontiguousarray(C, dtype=float)
    Qx, Qy, Fz = np.asarray(Qx, float), np.asarray(Qy, float), np.asarray(Fz, float)
    V0, V1 = np.asarray(V0, float), np.asarray(V1, float)
    okc = np.broadcast_to(ok, C.shape[1:])
    vert = bool(np.any(Fz != 0.0))
    ext = lambda c: c if inflow is None else np.asarray(inflow, float)          # noqa: E731
    dV = V1 - V0
    Va, Vb = V0 + dV, V0 + 0.5 * dV
    q0 = V0 * C
    if up_order:
        r = lambda c: rhs_up(c, Qx, Qy, Fz, okc, ext(c), open_mask, up_order)                   # noqa: E731
        fx = lambda c: faces_up_h(c, Qx, Qy, okc, ext(c), open_mask, up_order)[0]              # noqa: E731
    else:
        r = lambda c: rhs(c, Qx, Qy, Fz, okc, ext(c), open_mask, z_variant, vert)               # noqa: E731
        fx = lambda c: faces_h(c, Qx, Qy, okc, ext(c), open_mask, z_variant)[0]                # noqa: E731
    q1 = q0 + dt * r(C)
    c1 = _safe(q1, Va, C)
    q2 = 0.75 * q0 + 0.25 * (q1 + dt * r(c1))
    c2 = _safe(q2, Vb, C)
    q3 = q0 / 3.0 + 2.0 / 3.0 * (q2 + dt * r(c2))
    out = _safe(q3, V1, C)
    if flux_face_i is None:
        return out
    fxi = lambda c: fx(c)[..., flux_face_i]                                                     # noqa: E731
    return out, (fxi(C) + fxi(c1)) / 6.0 + 2.0 * fxi(c2) / 3.0


def advect_pc(C, V0, V1, Qx, Qy, Fz, ok, dt, inflow=None, open_mask=0, z_variant=True, flux_face_i=None,
              up_order=0):
    """
    One 'pc' step: the predictor ch = safe(V0 C + 0.5 dt r_c(C), V0 + 0.5 dV) with centred faces on x, y and z,
    then out = safe(V0 C + dt r(ch), V1) with the WENO (``z_variant``) or, with ``up_order`` 3 / 1, the UP faces;
    ``safe`` falls back to C.  Follows ``advect_numpy`` (tracers.py:303-330).  Returns out or (out, F(ch) at u-face
    column flux_face_i).
    """
    C = np.ascontiguousarray(C, dtype=float)
    Qx, Qy, Fz = np.asarray(Qx, float), np.asarray(Qy, float), np.asarray(Fz, float)
    V0, V1 = np.asarray(V0, float), np.asarray(V1, float)
    okc = np.broadcast_to(ok, C.shape[1:])
    ext = lambda c: c if inflow is None else np.asarray(inflow, float)          # noqa: E731
    dV = V1 - V0
    q0 = V0 * C
    ch = _safe(q0 + 0.5 * dt * rhs_up(C, Qx, Qy, Fz, okc, ext(C), open_mask, "c"), V0 + 0.5 * dV, C)
    if up_order:
        r = rhs_up(ch, Qx, Qy, Fz, okc, ext(ch), open_mask, up_order)
        fx = lambda: faces_up_h(ch, Qx, Qy, okc, ext(ch), open_mask, up_order)[0]           # noqa: E731
    else:
        r = rhs(ch, Qx, Qy, Fz, okc, ext(ch), open_mask, z_variant, bool(np.any(Fz != 0.0)))
        fx = lambda: faces_h(ch, Qx, Qy, okc, ext(ch), open_mask, z_variant)[0]            # noqa: E731
    out = _safe(q0 + dt * r, V1, C)
    return out if flux_face_i is None else (out, fx()[..., flux_face_i])


# ------------------------------------------------------------------------------------------------ periodic
TRACER_HALO = 13         # 3 stages x 3 cells + FCT 3 + 1 (core/model/periodic.py:31)


def pad(a, kind, halo, px, py):
    """Wrapped halo of ``halo`` cells/faces on the periodic axes (core/model/periodic.py:71-87)."""
    a = np.asarray(a)
    if px:
        face = kind in ("u", "q")
        n = a.shape[-1] - face
        a = np.take(a, np.arange(-halo, n + halo + face) % n, axis=-1)
    if py:
        face = kind in ("v", "q")
        n = a.shape[-2] - face
        a = np.take(a, np.arange(-halo, n + halo + face) % n, axis=-2)
    return a


def crop(a, halo, px, py):
    """Inverse of ``pad`` for cell arrays (core/model/periodic.py:89-95)."""
    if px:
        a = a[..., halo:a.shape[-1] - halo]
    if py:
        a = a[..., halo:a.shape[-2] - halo, :]
    return np.ascontiguousarray(a)


# ------------------------------------------------------------------------------------------------ diagnostics
def numerical_mixing(c, c2, V1, dt):
    """
    Numerical mixing of the advection of c, V1 (<c^2> - c^2) / dt per cell, with <c^2> the advected c^2
    (Burchard & Rennau 2008, eq. 24): ``VarianceBudget.advective`` (core/components/variance_budget.py:60-62) and
    ``mix_num`` (core/model/driver.py:905), same operation order.
    """
    return V1 * (c2 - c ** 2) / dt


# ------------------------------------------------------------------------------------------------ tracer diffusion
def diffuse_h_ale(T, Tw, h, act, cx, cy, okx, oky, V, kh, dt):
    """
    Explicit Laplacian diffusion along the layers, ALE form: the ALE branch of ``TracerDiffusion.diffuse_h_numpy``
    (core/components/tracers.py:450-481, with ``face_thickness``, core/vertical/ale.py:66-79), whose per-element
    arithmetic ``_diffuse_ale_kernel`` (core/dynamics/sigma_baroclinic.py:533-561) repeats.  T, Tw (..., n, ny, nx);
    h (n, ny, nx) zeroed on inactive cells; cx = L_u/dx_u (ny, nx+1), cy = L_v/dy_v (ny+1, nx); okx, oky face masks.
    """
    n, ny, nx = h.shape
    hu = np.zeros((n, ny, nx + 1))
    hv = np.zeros((n, ny + 1, nx))
    hu[..., 1:-1] = np.where((act[:, :-1] & act[:, 1:])[None], np.minimum(h[..., :-1], h[..., 1:]), 0.0)
    hv[:, 1:-1, :] = np.where((act[:-1, :] & act[1:, :])[None], np.minimum(h[:, :-1, :], h[:, 1:, :]), 0.0)
    lead = T.shape[:-3]
    Fx = np.zeros(lead + (n, ny, nx + 1))
    Fy = np.zeros(lead + (n, ny + 1, nx))
    Fx[..., 1:-1] = -kh * cx[None, :, 1:-1] * hu[..., 1:-1] * (Tw[..., 1:] - Tw[..., :-1])
    Fy[..., 1:-1, :] = -kh * cy[None, 1:-1, :] * hv[:, 1:-1, :] * (Tw[..., 1:, :] - Tw[..., :-1, :])
    Fx = np.where(okx[None], Fx, 0.0)
    Fy = np.where(oky[None], Fy, 0.0)
    div = Fx[..., 1:] - Fx[..., :-1] + Fy[..., 1:, :] - Fy[..., :-1, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(V > 0, T - dt * div / np.where(V > 0, V, 1.0), T)


def diffuse_h_sigma(T, Tw, ds, gx, gy, okx, oky, V, kh, dt):
    """
    Explicit Laplacian diffusion along sigma surfaces: the sigma branch of ``TracerDiffusion.diffuse_h_numpy``
    (core/components/tracers.py:450-481, h is None): F_x = -K_h G_u ds_k (Tw_i - Tw_i-1) on the interior faces with
    G_u = L_u nan_to_num(D_u) / dx_u (gx, (ny, nx+1)) and likewise G_v (gy), masked by okx / oky.  T, Tw
    (..., n, ny, nx); ds (n,); V (n, ny, nx).
    """
    n = ds.shape[0]
    ny, nx = gx.shape[0], gx.shape[1] - 1
    d = np.asarray(ds, float)[:, None, None]
    lead = T.shape[:-3]
    Fx = np.zeros(lead + (n, ny, nx + 1))
    Fy = np.zeros(lead + (n, ny + 1, nx))
    Fx[..., 1:-1] = -kh * gx[None, :, 1:-1] * d * (Tw[..., 1:] - Tw[..., :-1])
    Fy[..., 1:-1, :] = -kh * gy[None, 1:-1, :] * d * (Tw[..., 1:, :] - Tw[..., :-1, :])
    Fx = np.where(okx[None], Fx, 0.0)
    Fy = np.where(oky[None], Fy, 0.0)
    div = Fx[..., 1:] - Fx[..., :-1] + Fy[..., 1:, :] - Fy[..., :-1, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(V > 0, T - dt * div / np.where(V > 0, V, 1.0), T)


def vertical_implicit(u, dz, nu, r_bed, dt, tau_top=None):
    """
    Backward-Euler vertical diffusion per column, Thomas algorithm: ``_vertical_implicit_numpy`` and its compiled
    twin ``_thomas_columns`` (core/dynamics/sigma_wd.py:57-88, 112-155; same arithmetic), with the compiled entry's
    zero surface flux when ``tau_top`` is None (``vertical_implicit`` :91-99).  u, dz (n, M); nu (n+1, M);
    r_bed (M,) on layer 0 or (n, M).  A leading stack axis of u is solved field by field.
    """
    u = np.asarray(u, float)
    if u.ndim == 3:
        return np.stack([vertical_implicit(x, dz, nu, r_bed, dt, tau_top) for x in u])
    n, M = u.shape
    r = np.asarray(r_bed, float)
    rates = np.ascontiguousarray(np.broadcast_to(r, (n, M))) if r.ndim == 2 else np.zeros((n, M))
    if r.ndim != 2:
        rates[0] = np.broadcast_to(r, (M,))
    h = 0.5 * (dz[1:] + dz[:-1])
    c = dt * nu[1:-1] / h
    lower = np.zeros_like(u)
    upper = np.zeros_like(u)
    lower[1:] = -c
    upper[:-1] = -c
    diag = dz.copy()
    diag[1:] += c
    diag[:-1] += c
    diag += dt * rates
    rhs = dz * u
    rhs[-1] = rhs[-1] + dt * (np.zeros(M) if tau_top is None else tau_top)
    cp = np.zeros_like(u)
    dp = np.zeros_like(u)
    cp[0] = upper[0] / diag[0]
    dp[0] = rhs[0] / diag[0]
    for k in range(1, n):
        m = diag[k] - lower[k] * cp[k - 1]
        cp[k] = upper[k] / m
        dp[k] = (rhs[k] - lower[k] * dp[k - 1]) / m
    out = np.empty_like(u)
    out[-1] = dp[-1]
    for k in range(n - 2, -1, -1):
        out[k] = dp[k] - cp[k] * out[k + 1]
    return out


# ------------------------------------------------------------------------------------------------ ALE layer transport
H_PPM_MIN = 2e-10        # positive-definite floor of the PPM thickness (core/vertical/ale.py:298)


def face_transport_thickness(h, H, active, clip):
    """
    Transport thickness at the u-faces, (q_e, q_w) each (n, ny, nx+1): the upwind layer of the west (q_e) and east
    (q_w) cell, cut at the shallower bed when ``clip``; 0 on faces without water on both sides.  Follows the
    compiled ``_ftt_kernel`` (core/vertical/ale.py:168-199): NaN thickness or depth read as 0, running bed-up sums.
    For the v-faces call it on swapped axes.
    """
    h = np.asarray(h, float)
    n, ny, nx = h.shape
    act = np.asarray(active, bool)
    f = np.arange(nx + 1)
    iw, ie = np.maximum(f - 1, 0), np.minimum(f, nx - 1)
    open_ = act[:, iw] & act[:, ie]
    hn = np.where(np.isnan(h), 0.0, h)
    hw, he = hn[..., iw], hn[..., ie]
    if clip:
        Hn = np.asarray(H, float)
        Hn = np.where(np.isnan(Hn), 0.0, Hn)
        Hw, He = Hn[:, iw], Hn[:, ie]
        z_bed = np.maximum(-Hw, -He)
        cw = np.zeros((ny, nx + 1))
        ce = np.zeros((ny, nx + 1))
        qe = np.empty((n, ny, nx + 1))
        qw = np.empty((n, ny, nx + 1))
        for k in range(n):
            cw = cw + hw[k]
            ce = ce + he[k]
            qe[k] = np.minimum(hw[k], np.maximum(0.0, (-Hw + cw) - z_bed))
            qw[k] = np.minimum(he[k], np.maximum(0.0, (-He + ce) - z_bed))
    else:
        qe, qw = hw, he
    return np.where(open_[None], qe, 0.0), np.where(open_[None], qw, 0.0)


def transport_velocity(u_new, u_old, last_wet, w, L, D, theta, theta_bc=None):
    """
    Transport velocity and free-surface volume flux of the ALE step, (u^θ, Q_bt):
    u^θ_k = θ_bc u_k' + (1 - θ_bc) u_k + (θ - θ_bc)(Σ w u' - Σ w u), u = 0 on faces dry in the last step,
    Q_bt = L D Σ_k w_k u^θ_k.  Follows ``_transport_velocity_kernel`` (core/vertical/ale.py:789-808), sums from
    the bed up.  u (n, a, b); last_wet, L, D (a, b).
    """
    tb = theta if theta_bc is None else theta_bc
    un = np.where(np.asarray(last_wet, bool)[None], u_old, 0.0)
    n = u_new.shape[0]
    m1 = np.zeros(u_new.shape[1:])
    m0 = np.zeros(u_new.shape[1:])
    for k in range(n):
        m1 = m1 + w[k] * u_new[k]
        m0 = m0 + w[k] * un[k]
    d = (theta - tb) * (m1 - m0)
    ut = tb * u_new + (1 - tb) * un + d[None]
    s = np.zeros(u_new.shape[1:])
    for k in range(n):
        s = s + w[k] * ut[k]
    return ut, L * D * s


def ppm_edges(h, active, h_min=H_PPM_MIN):
    """
    West and east edge thickness (h_w, h_e) of every cell along x: limited slopes (zero next to land and at the
    domain edge), edges from the neighbour means and slopes, then the positive-definite limiter that scales a
    parabola whose minimum lies inside the cell, or flattens a cell with h <= h_min.  Follows the compiled
    ``_ppm_edges_kernel`` (core/vertical/ale.py:578-623) term for term, including its (3 dh) dh; the input is
    read as ``ppm_transport_x`` passes it (NaN -> 0, inactive cells 0).
    """
    act = np.asarray(active, bool)
    h = np.where(act[None], np.nan_to_num(np.asarray(h, float)), 0.0)
    n, ny, nx = h.shape
    m = act.astype(float)
    mw = np.zeros((ny, nx))
    me = np.zeros((ny, nx))
    mw[:, 1:] = m[:, :-1]
    me[:, :-1] = m[:, 1:]
    hm1 = h.copy()
    hp1 = h.copy()
    hm1[..., 1:] = mw[None, :, 1:] * h[..., :-1] + (1.0 - mw[None, :, 1:]) * h[..., 1:]
    hp1[..., :-1] = me[None, :, :-1] * h[..., 1:] + (1.0 - me[None, :, :-1]) * h[..., :-1]
    s = 0.5 * (hp1 - hm1)
    dmx = np.maximum(np.maximum(hp1, hm1), h) - h
    dmn = h - np.minimum(np.minimum(hp1, hm1), h)
    a = np.minimum(np.abs(s), 2.0 * np.minimum(dmx, dmn))
    s = np.where(s > 0.0, a, np.where(s < 0.0, -a, 0.0))
    slp = np.where((mw * me * m > 0.0)[None], s, 0.0)
    sw = np.zeros_like(slp)
    se = np.zeros_like(slp)
    sw[..., 1:] = slp[..., :-1]
    se[..., :-1] = slp[..., 1:]
    hw = 0.5 * (hm1 + h) + (sw - slp) / 6.0
    he = 0.5 * (hp1 + h) + (slp - se) / 6.0
    curv = 3.0 * ((hw + he) - 2.0 * h)
    dh = he - hw
    inside = (curv > 0.0) & (np.abs(dh) < curv)
    flat = inside & (h <= h_min)
    num = 12.0 * curv * (h - h_min)
    den = curv * curv + 3.0 * dh * dh
    with np.errstate(divide="ignore", invalid="ignore"):
        sc = np.where(inside & (num < den), num / den, 1.0)
    out_w = np.where(flat, h, h + sc * (hw - h))
    out_e = np.where(flat, h, h + sc * (he - h))
    return out_w, out_e


def _ppm_face_flux(u, h, h_w, h_e, dx_t, q_e, q_w, L, dt):
    """(F, dF/du) of every layer at every u-face for face velocities u (n, ny, nx+1): the swept-volume mean
    thickness of the donor cell's parabola and its derivative factor, cut by the transport thickness where that
    is below the donor layer.  Follows ``_face_flux`` (core/vertical/ale.py:626-648)."""
    nx = h.shape[-1]
    f = np.arange(nx + 1)
    iw, ie = np.maximum(f - 1, 0), np.minimum(f, nx - 1)
    pos = u >= 0.0
    dxd = np.where(pos, dx_t[None][..., iw], dx_t[None][..., ie])
    c = np.minimum(np.abs(u) * dt / dxd, 1.0)
    hu = np.where(pos, h[..., iw], h[..., ie])
    ww = np.where(pos, h_w[..., iw], h_w[..., ie])
    ee = np.where(pos, h_e[..., iw], h_e[..., ie])
    curv3 = (ww + ee) - 2.0 * hu
    dh = np.where(pos, ww - ee, ee - ww)
    edge = np.where(pos, ee, ww)
    q = np.where(pos, q_e, q_w)
    mean = edge + c * (0.5 * dh + curv3 * (c - 1.5))
    marg = edge + c * (dh + 3.0 * curv3 * (c - 1.0))
    clip = q < hu
    mean = np.where(clip, np.minimum(mean, q), mean)
    marg = np.where(clip, np.minimum(marg, q), marg)
    return L[None] * u * np.maximum(mean, 0.0), L[None] * np.maximum(marg, 0.0)


def ppm_transport_x(h, ut, L, Q_bt, q_e, q_w, active, dt, dx_t, max_iter=20, return_iterations=False):
    """
    Layer volume fluxes at the u-faces with PPM thickness and a uniform velocity correction du per face column,
    found by Newton's method with bracketing bisection so that Σ_k F_k(u_k + du) = Q_bt; every column iterates
    until all columns have converged (|err| <= 1e-12 max(|Q_bt|, 1)) or ``max_iter`` passes, then the remaining
    residual R = Q_bt - Σ F is shared over the layers in proportion to the transport thickness on R's side.
    Follows the compiled ``_ppm_transport_kernel`` (core/vertical/ale.py:651-740) and its wrapper (:743-764); sums
    run from the bed up.  For the v-faces call it on swapped axes.
    """
    act = np.asarray(active, bool)
    hc = np.where(act[None], np.nan_to_num(np.asarray(h, float)), 0.0)
    n, ny, nx = hc.shape
    h_w, h_e = ppm_edges(hc, act)
    ut = np.asarray(ut, float)
    L = np.asarray(L, float)
    Q_bt = np.asarray(Q_bt, float)
    q_e = np.asarray(q_e, float)
    q_w = np.asarray(q_w, float)
    dx_t = np.asarray(dx_t, float)
    f = np.arange(nx + 1)
    iw, ie = np.maximum(f - 1, 0), np.minimum(f, nx - 1)
    open_ = act[:, iw] & act[:, ie]
    limw, lime = 0.5 * dx_t[:, iw] / dt, 0.5 * dx_t[:, ie] / dt
    mnw = np.full((ny, nx + 1), np.inf)
    mxe = np.full((ny, nx + 1), -np.inf)
    for k in range(n):
        mnw = np.minimum(mnw, limw - ut[k])
        mxe = np.maximum(mxe, -(lime + ut[k]))
    hi = np.maximum(np.minimum(limw, mnw), 0.0)
    lo = np.minimum(np.maximum(-lime, mxe), 0.0)

    def evaluate(u):
        F, dF = _ppm_face_flux(u, hc, h_w, h_e, dx_t, q_e, q_w, L, dt)
        s = np.zeros((ny, nx + 1))
        sd = np.zeros((ny, nx + 1))
        for k in range(n):
            s = s + F[k]
            sd = sd + dF[k]
        return F, np.where(open_, s, 0.0) - Q_bt, np.where(open_, sd, 0.0)

    du = np.zeros((ny, nx + 1))
    F, err, dsum = evaluate(ut)
    iters = 0
    with np.errstate(divide="ignore", invalid="ignore"):
        for _ in range(max_iter):
            hi = np.where(err > 0.0, du, hi)
            lo = np.where(err < 0.0, du, lo)
            step = np.where(dsum > 0.0, -err / dsum, 0.0)
            new = du + step
            new = np.where((step > 0.0) & (new >= hi), 0.5 * (du + hi), new)
            new = np.where((step < 0.0) & (new <= lo), 0.5 * (du + lo), new)
            du = new
            F, err, dsum = evaluate(ut + du[None])
            iters += 1
            if not (np.abs(err) > 1e-12 * np.maximum(np.abs(Q_bt), 1.0)).any():
                break
    s = np.zeros((ny, nx + 1))
    te = np.zeros((ny, nx + 1))
    tw = np.zeros((ny, nx + 1))
    for k in range(n):
        s = s + F[k]
        te = te + q_e[k]
        tw = tw + q_w[k]
    R = Q_bt - s
    use_e = R >= 0.0
    tot = np.where(use_e, te, tw)
    use_e = np.where(tot <= 0.0, ~use_e, use_e)
    tot = np.where(use_e, te, tw)
    den = np.where(tot > 0.0, tot, 1.0)
    out = np.where(open_[None], F + R[None] * np.where(use_e[None], q_e, q_w) / den[None], 0.0)
    return (out, iters) if return_iterations else out


def lagrangian_thickness(h, divq, dt, area, tol=1e-9):
    """
    Layer thickness after the Lagrangian step, h - Δt ∇·Q / A (A = 1 where the area is 0); values in [-tol, 0) are
    set to 0 and a layer emptied beyond that raises with its (k, j, i).  Follows ``lagrangian_thickness``
    (core/vertical/ale.py:278-295).
    """
    out = h - dt * divq / np.where(area > 0, area, 1.0)[None]
    bad = out < -tol
    if bad.any():
        k, j, i = (int(a[0]) for a in np.nonzero(bad))
        raise RuntimeError(f"layer {k} of cell (j={j}, i={i}) emptied in the Lagrangian step: "
                           f"h = {out[k, j, i]:.3e} m (reduce dt)")
    return np.maximum(out, 0.0)


def outflow(Qx, Qy):
    """Outgoing volume flux of every cell-layer over its four faces (core/vertical/ale.py:301-304)."""
    return (np.maximum(Qx[..., 1:], 0.0) - np.minimum(Qx[..., :-1], 0.0)
            + np.maximum(Qy[:, 1:, :], 0.0) - np.minimum(Qy[:, :-1, :], 0.0))


def _sides(a, ax, wrap, fill):
    """Cell values on the low and high side of every face along axis ``ax``: (low, high), each one face longer
    than ``a``; the edge faces take ``fill`` or, on a periodic axis, the cell across the wrap."""
    first, last = np.take(a, [0], axis=ax), np.take(a, [-1], axis=ax)
    pad = np.full_like(first, fill)
    low = np.concatenate([last if wrap else pad, a], axis=ax)
    high = np.concatenate([a, first if wrap else pad], axis=ax)
    return low, high


def _sum_k(a):
    s = np.zeros(a.shape[1:])
    for k in range(a.shape[0]):
        s = s + a[k]
    return s


def _give_back(Q, delta, room, ax, full, wrap):
    """The net cut of each face back on its layers, in place: to same-way layers by donor room, then off
    counter-flowing layers by |Q| (never past 0), then on the donor's layers by content.  Follows ``_give_back``
    (core/vertical/ale.py:399-437); raises as it does when a donor column holds no water."""
    room_lo, room_hi = _sides(room, ax, wrap, 0.0)
    room_lo, room_hi = np.maximum(room_lo, 0.0), np.maximum(room_hi, 0.0)
    s = np.sign(delta)[None]
    same = (np.sign(Q) == s) & (Q != 0.0) & (s != 0.0)
    rm = np.where(same, np.where(s > 0, room_lo, room_hi), 0.0)
    tot = _sum_k(rm)
    need = np.abs(delta)
    give = np.minimum(need, tot)
    with np.errstate(divide="ignore", invalid="ignore"):
        Q += s * rm * np.where(tot > 0, give / np.where(tot > 0, tot, 1.0), 0.0)[None]
        rest = need - give
        opp = (np.sign(Q) == -s) & (s != 0.0)
        cap = _sum_k(np.where(opp, np.abs(Q), 0.0))
        frac = np.where(cap > 0, np.minimum(rest, cap) / np.where(cap > 0, cap, 1.0), 0.0)[None]
        Q -= np.where(opp, Q * frac, 0.0)
        rest = rest - np.minimum(rest, cap)
        full_lo, full_hi = _sides(full, ax, wrap, 0.0)
        fm = np.where(s > 0, full_lo, full_hi) * (rest > 0)[None]
        tot = _sum_k(fm)
        if ((rest > 0) & (tot <= 0)).any():
            idx = tuple(int(a[0]) for a in np.nonzero((rest > 0) & (tot <= 0)))
            raise RuntimeError(f"limit_outflow: face {idx} must carry {float(rest[idx]):.3g} m³/s out of a donor "
                               f"column that holds no water — barotropic CFL")
        Q += s * fm * np.where(tot > 0, rest / np.where(tot > 0, tot, 1.0), 0.0)[None]


def limit_outflow(h, Qx, Qy, area, dt, active, h_floor=H_PPM_MIN, max_pass=50, periodic=(False, False)):
    """
    Outflow limiter of the Lagrangian step: a cell-layer whose outgoing fluxes would take more than
    (h - h_floor)+ A in Δt has them scaled down to that, and each face's net cut goes back to the other layers of
    the face (``_give_back``), so Σ_k Q_k is unchanged; repeated until no cell-layer exceeds (<= ``max_pass``).
    Returns (Qx, Qy, n_limited, status) with status "ok", "noop" (inputs returned) or "accepted" (still exceeding
    after max_pass but h >= -h_floor everywhere); raises where the reference raises.  Follows ``limit_outflow``
    (core/vertical/ale.py:331-396), sums over the layers from the bed up.
    """
    act = np.asarray(active, bool)
    A = np.where(act, area, 0.0)[None]
    room0 = np.maximum(h - h_floor, 0.0) * A / dt
    O = outflow(Qx, Qy)
    over = act[None] & (O > room0)
    if not over.any():
        return Qx, Qy, 0, "noop"
    Qx, Qy, n_lim = Qx.copy(), Qy.copy(), 0
    for _ in range(max_pass):
        n_lim += int(over.sum())
        with np.errstate(divide="ignore", invalid="ignore"):
            r = np.where(over, room0 / np.where(O > 0, O, 1.0), 1.0)
        for Q, ax, wrap in ((Qx, 2, periodic[0]), (Qy, 1, periodic[1])):
            r_lo, r_hi = _sides(r, ax, wrap, 1.0)
            new = np.where(Q > 0, Q * r_lo, np.where(Q < 0, Q * r_hi, Q))
            delta = _sum_k(Q - new)
            Q[...] = new
            _give_back(Q, delta, room0 - outflow(Qx, Qy), ax, room0, wrap)
        O = outflow(Qx, Qy)
        over = act[None] & (O > room0 * (1.0 + 1e-12))
        if not over.any():
            return Qx, Qy, n_lim, "ok"
    div = Qx[..., 1:] - Qx[..., :-1] + Qy[:, 1:] - Qy[:, :-1]
    with np.errstate(divide="ignore", invalid="ignore"):
        h_dag = np.where(act[None], h - dt * div / np.where(A > 0, A, 1.0), 0.0)
    if h_dag.min() >= -h_floor:
        return Qx, Qy, n_lim, "accepted"
    k, j, i = (int(a[0]) for a in np.nonzero(h_dag == h_dag.min()))
    raise RuntimeError(f"limit_outflow: layer {k} of cell (j={j}, i={i}) would reach h = {h_dag.min():.3g} m after "
                       f"{max_pass} passes — barotropic CFL")


def transport_thickness(h, H, active, clip=True, periodic=(False, False)):
    """
    The vertical coordinate's transport thicknesses ((q_e, q_w) at the u-faces, (q_n, q_s) at the v-faces in
    their natural (n, ny+1, nx) layout): ``face_transport_thickness`` on both axes, on periodic axes padded by one
    wrapped cell and cropped (``ALECoordinate.transport_thickness``, core/components/vertical_coordinate.py:125-140).
    """
    px, py = periodic
    act = np.asarray(active, bool)
    if px or py:
        hp, Hp, ap = pad(h, "c", 1, px, py), pad(H, "c", 1, px, py), pad(act, "c", 1, px, py)
        x = tuple(crop(q, 1, px, py) for q in face_transport_thickness(hp, Hp, ap, clip))
        y = tuple(crop(np.ascontiguousarray(q.swapaxes(1, 2)), 1, px, py)
                  for q in face_transport_thickness(hp.swapaxes(1, 2), Hp.T, ap.T, clip))
    else:
        x = face_transport_thickness(h, H, act, clip)
        y = tuple(np.ascontiguousarray(q.swapaxes(1, 2))
                  for q in face_transport_thickness(h.swapaxes(1, 2), H.T, act.T, clip))
    return x, y


def upwind_transport_x(q_e, q_w, ut, L, Q_bt):
    """
    Layer volume fluxes through the u-faces with the donor-cell thickness, corrected to the free-surface flux:
    Q_k = L u_k q_k(u_k) + R q_k(R) / sum_j q_j(R) with R = Q_bt - sum_k L u_k q_k(u_k), q(s) the transport
    thickness on the side the water comes from (the other side's where that column is empty).  Follows
    ``upwind_transport_x`` (core/vertical/ale.py:233-251) term for term.  For the v-faces call it on swapped axes.
    """
    F = L[None] * ut * np.where(ut >= 0, q_e, q_w)
    R = Q_bt - F.sum(axis=0)
    qR = np.where((R >= 0)[None], q_e, q_w)
    alt = np.where((R >= 0)[None], q_w, q_e)
    qR = np.where((qR.sum(axis=0) > 0)[None], qR, alt)
    tot = qR.sum(axis=0)
    return F + R[None] * qR / np.where(tot > 0, tot, 1.0)[None]


def open_face_transport(h_int, u_new, u_old, Q_bt, L, theta, theta_bc=None):
    """
    Layer volume fluxes through the faces of one open edge: Q_k = (h_k / sum h) Q_bt + L h_k (tb u'_new + (1 - tb)
    u'_old) with the baroclinic parts u' = u - sum (h/sum h) u, the interior cell's thicknesses h_int (N, m); 0 on a
    column without water.  Follows ``open_face_transport`` (core/vertical/ale.py:841-861) term for term.
    """
    tot = h_int.sum(axis=0)
    wet = tot > 0
    w = np.where(wet[None], h_int / np.where(wet, tot, 1.0)[None], 0.0)
    pn = u_new - (w * u_new).sum(axis=0)[None]
    po = u_old - (w * u_old).sum(axis=0)[None]
    tb = theta if theta_bc is None else theta_bc
    Q = w * Q_bt[None] + L[None] * h_int * (tb * pn + (1.0 - tb) * po)
    return np.where(wet[None], Q, 0.0)


def open_face_transports(Qx, Qy, h, un, vn, u_new, v_new, open_edges, open_Du, open_Qu, open_Dv, open_Qv, Lu, Lv,
                         theta, theta_bc=None):
    """
    In place on Qx, Qy: the open edges' faces where the free-surface face depth is positive take
    ``open_face_transport`` of the interior cells (``OpenBoundary.open_face_transports``,
    core/components/open_boundary.py:334-351); un, vn the old velocities, u_new, v_new the new ones.
    """
    for edge, Q, D, Qo, L, hc, vo, vn1 in (
            ("west", Qx[..., 0], open_Du[:, 0], open_Qu[:, 0], Lu[:, 0], h[..., 0], un[..., 0], u_new[..., 0]),
            ("east", Qx[..., -1], open_Du[:, -1], open_Qu[:, -1], Lu[:, -1], h[..., -1], un[..., -1],
             u_new[..., -1]),
            ("south", Qy[:, 0, :], open_Dv[0, :], open_Qv[0, :], Lv[0, :], h[:, 0, :], vn[:, 0, :], v_new[:, 0, :]),
            ("north", Qy[:, -1, :], open_Dv[-1, :], open_Qv[-1, :], Lv[-1, :], h[:, -1, :], vn[:, -1, :],
             v_new[:, -1, :])):
        if edge in open_edges:
            Q[...] = np.where((D > 0)[None], open_face_transport(hc, vn1, vo, Qo, L, theta, theta_bc), Q)
    return Qx, Qy


def ale_layer_transport(h, H, active, area, dx_t, dy_t, Lu, Lv, u_new, u_old, v_new, v_old, wet_u, wet_v, wu, wv,
                        Du, Dv, dt, theta, theta_bc=None, clip=True, max_iter=20, periodic=(False, False),
                        thickness_flux="ppm", fw=None, open_=(), open_Du=None, open_Qu=None, open_Dv=None,
                        open_Qv=None):
    """
    One ALE layer-transport step with PPM thickness: transport velocities, transport thicknesses (the vertical
    coordinate's, clipped at the shallower bed for z*), PPM fluxes of both directions (the v-faces on swapped axes),
    the outflow limiter, the Lagrangian thickness and the layer volumes.  Follows ``ALELayerTransport.__call__``
    (core/components/layer_transport.py:111-156), with
    ``transport_thickness`` (core/components/vertical_coordinate.py:125-140) and ``layer_volumes``
    (core/dynamics/sigma_baroclinic.py:1812-1816).  On periodic axes the PPM fluxes run on inputs padded by 4
    wrapped cells and are cropped (layer_transport.py:122-129, core/model/periodic.py:71-103).  Returns
    dict(Qx, Qy, V0, V1, ok, h_dagger, n_limited).  ``thickness_flux`` "upwind": the donor-cell fluxes of
    ``upwind_transport_x`` instead of PPM (layer_transport.py:131-133), with no padding.  With the freshwater flux
    ``fw`` (ny, nx) [m/s] also h_fw (h† with dt fw on the top layer, ``freshwater_volumes`` with the floor 1e-6 m)
    and V_fw = h_fw A (layer_transport.py:148-155).  ``open_`` edges take ``open_face_transports`` of the open
    flux fields before the limiter (layer_transport.py:134-135).
    """
    act = np.asarray(active, bool)
    px, py = periodic
    utx, Qbx = transport_velocity(u_new, u_old, wet_u, wu, Lu, Du, theta, theta_bc)
    uty, Qby = transport_velocity(v_new, v_old, wet_v, wv, Lv, Dv, theta, theta_bc)
    (qe, qw), (qn_nat, qs_nat) = transport_thickness(h, H, act, clip, periodic)
    qn, qs = qn_nat.swapaxes(1, 2), qs_nat.swapaxes(1, 2)
    sw = lambda a: a.swapaxes(1, 2)                                  # noqa: E731
    if thickness_flux == "upwind":
        Qx = upwind_transport_x(qe, qw, utx, Lu, Qbx)
        Qy = upwind_transport_x(qn, qs, sw(uty), Lv.T, Qby.T).swapaxes(1, 2)
    elif px or py:
        P = lambda a, k: pad(a, k, 4, px, py)                        # noqa: E731
        hp, ap = P(h, "c"), P(act, "c")
        Qx = crop(ppm_transport_x(hp, P(utx, "u"), P(Lu, "u"), P(Qbx, "u"), P(qe, "u"), P(qw, "u"), ap, dt,
                                  P(dx_t, "c"), max_iter), 4, px, py)
        Qy = crop(np.ascontiguousarray(sw(ppm_transport_x(sw(hp), sw(P(uty, "v")), P(Lv, "v").T, P(Qby, "v").T,
                                                          sw(P(qn_nat, "v")), sw(P(qs_nat, "v")), ap.T, dt,
                                                          P(dy_t, "c").T, max_iter))), 4, px, py)
    else:
        Qx = ppm_transport_x(h, utx, Lu, Qbx, qe, qw, act, dt, dx_t, max_iter)
        Qy = ppm_transport_x(sw(h), sw(uty), Lv.T, Qby.T, qn, qs, act.T, dt, dy_t.T, max_iter).swapaxes(1, 2)
    if open_:
        Qx, Qy = np.ascontiguousarray(Qx), np.ascontiguousarray(Qy)
        open_face_transports(Qx, Qy, h, u_old, v_old, u_new, v_new, open_, open_Du, open_Qu, open_Dv, open_Qv, Lu,
                             Lv, theta, theta_bc)
    Qx, Qy, n_lim, _ = limit_outflow(h, np.ascontiguousarray(Qx), np.ascontiguousarray(Qy), area, dt, act,
                                     periodic=(px, py))
    divh = Qx[..., 1:] - Qx[..., :-1] + Qy[:, 1:, :] - Qy[:, :-1, :]
    h_lag = lagrangian_thickness(h, divh, dt, area)
    V0, V1 = h * area[None], h_lag * area[None]
    h_fw = V_fw = None
    if fw is not None:
        h_fw, _ = freshwater_volumes(h_lag, dt * fw, 1e-6)
        V_fw = h_fw * area[None]
    return dict(Qx=np.ascontiguousarray(Qx), Qy=np.ascontiguousarray(Qy), V0=V0, V1=V1, ok=act[None] & (V1 > 0),
                h_dagger=h_lag, n_limited=n_lim, qe=qe, qw=qw, qn=np.ascontiguousarray(qn.swapaxes(1, 2)),
                qs=np.ascontiguousarray(qs.swapaxes(1, 2)), utx=utx, uty=uty, Qbx=Qbx, Qby=Qby, h_fw=h_fw, V_fw=V_fw)


# ------------------------------------------------------------------------------------------------ layer density
# Wright (1997), Table 1, second column (full-range fit), the coefficients MOM6's WRIGHT_FULL uses; the same
# decimal values as physics/eos/wright_full.py:63-77.
WRIGHT = dict(a0=7.133718e-4, a1=2.724670e-7, a2=-1.646582e-7, b0=5.613770e8, b1=3.600337e6, b2=-3.727194e4,
              b3=1.660557e2, b4=6.844158e5, b5=-8.389457e3, c0=1.609893e5, c1=8.427815e2, c2=-6.931554,
              c3=3.869318e-2, c4=-1.664201e2, c5=-2.765195)


def linear_density(T, S=None, rho0=1000.0, t_coef=0.2, t_ref=0.0, s_coef=0.0, s_ref=35.0):
    """Linear EOS rho0 - t_coef (T - t_ref) [+ s_coef (S - s_ref)] [kg m^-3] (LinearEOS.density,
    core/components/eos.py:62-69; Vallis 2006 eq. 1.156)."""
    if S is None:
        return rho0 - t_coef * (T - t_ref)
    return rho0 - t_coef * (T - t_ref) + s_coef * (S - s_ref)


def wright_density(T, S, p):
    """Wright (1997) full-range in situ density rho(theta, S, p) [kg m^-3], p the sea pressure [Pa]
    (density_elem, physics/eos/wright_full.py:80-89, MOM6 WRIGHT_FULL parenthesization)."""
    w = WRIGHT
    al0 = w["a0"] + (w["a1"] * T + w["a2"] * S)
    p0 = w["b0"] + (w["b4"] * S + T * (w["b1"] + (T * (w["b2"] + w["b3"] * T) + w["b5"] * S)))
    lam = w["c0"] + (w["c4"] * S + T * (w["c1"] + (T * (w["c2"] + w["c3"] * T) + w["c5"] * S)))
    return (p + p0) / (lam + al0 * (p + p0))


# ------------------------------------------------------------------------------------------------ freshwater
def freshwater_volumes(V, dV, floor):
    """
    The freshwater volume dV (ny, nx) added to the top layer (k = N-1) of V (N, ny, nx); a removal that would take
    the top layer below ``floor`` is taken from the layers below, top-down, each first kept at the floor, then
    down to zero.  Returns (V_new, n_limited).  Follows ``freshwater_volumes`` (core/components/freshwater.py:31-58).
    """
    out = V.copy()
    out[-1] += dV
    floor = np.broadcast_to(np.asarray(floor, float), dV.shape)
    spill = (dV < 0) & (out[-1] < floor)
    if not spill.any():
        return out, 0
    js, is_ = np.nonzero(spill)
    col = V[:, js, is_].copy()
    deficit = -dV[js, is_]
    fl = floor[js, is_]
    for keep in (fl, np.zeros_like(fl)):
        for k in range(V.shape[0] - 1, -1, -1):
            take = np.minimum(deficit, np.maximum(col[k] - keep, 0.0))
            col[k] -= take
            deficit = deficit - take
    out[:, js, is_] = col
    return out, int(spill.sum())


def apply_freshwater(stack, layout, V_adv, V_fw, fw_area_dt, dt):
    """
    The freshwater substep on the tracer stack (in place, also returned) with (n_limited, var_surf): the members
    other than T and T2 keep their content (c V_adv / V_fw where the column has freshwater); a column whose volume
    changed below the top layer shares sum V_adv c / sum V_fw over the touched layers.  Follows
    ``apply_freshwater`` (core/components/freshwater.py:61-122).
    """
    cols = fw_area_dt != 0
    ratio = np.ones_like(V_adv)
    np.divide(V_adv, V_fw, out=ratio, where=cols[None] & (V_fw > 0))
    var_surf = None
    if "S" in layout:
        S_old = stack[layout["S"]].copy()
    salt_like = [i for i in range(layout["n"]) if i not in (layout["T"], layout.get("T2"))]
    spill = (V_adv[:-1] != V_fw[:-1]).any(axis=0)
    if spill.any():
        changed = V_adv != V_fw
        kmin = np.argmax(changed, axis=0)
        below = np.arange(V_adv.shape[0])[:, None, None] >= kmin[None]
        touched = below & spill[None]
        Vf_sum = np.where(touched, V_fw, 0.0).sum(axis=0)
        has_water = Vf_sum > 0
    n_limited = int(np.count_nonzero(spill))
    for i in salt_like:
        c = np.where(cols[None], stack[i] * ratio, stack[i])
        if spill.any():
            cbar = np.divide(np.where(touched, V_adv * stack[i], 0.0).sum(axis=0), Vf_sum,
                             out=np.zeros_like(Vf_sum), where=has_water)
            c = np.where(touched & has_water[None], cbar[None], c)
        stack[i] = c
    if "S" in layout:
        S_new = stack[layout["S"]]
        var_surf = np.where(cols[None], (V_fw * S_new**2 - V_adv * S_old**2) / dt, 0.0)
    return stack, n_limited, var_surf
