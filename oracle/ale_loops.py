"""
Scalar-loop restatement of the ALE layer-transport kernels (test infrastructure only; see ``oracle/__init__``).

One cell or face column at a time, in the order of the reference's compiled loops, with Python floats (IEEE FP64,
every operation rounded on its own).  It is a second, independent form of ``numpy_ref``'s vectorised restatement:
the tests check that the two agree bit for bit on small inputs, which catches index and padding slips that a
vectorised form can make.  Only for small inputs (pure Python).
"""
from __future__ import annotations

import math

import numpy as np

H_PPM_MIN = 2e-10


def _nan0(x: float) -> float:
    """numpy.nan_to_num for one value (NaN -> 0, +-inf -> +-largest finite)."""
    if x != x:
        return 0.0
    if x == math.inf:
        return np.finfo(float).max
    if x == -math.inf:
        return -np.finfo(float).max
    return x


def ppm_edges(h, active, h_min=H_PPM_MIN):
    """Edges of ``_ppm_edges_kernel`` (core/vertical/ale.py:578-623), one row of one layer at a time."""
    h = np.asarray(h, float)
    n, ny, nx = h.shape
    hc = [[[(_nan0(float(h[k, j, i])) if active[j, i] else 0.0) for i in range(nx)] for j in range(ny)]
          for k in range(n)]
    hw_out = np.empty((n, ny, nx))
    he_out = np.empty((n, ny, nx))
    for j in range(ny):
        for k in range(n):
            row = hc[k][j]
            slp = [0.0] * nx
            for i in range(nx):
                m = 1.0 if active[j, i] else 0.0
                mw = (1.0 if active[j, i - 1] else 0.0) if i > 0 else 0.0
                me = (1.0 if active[j, i + 1] else 0.0) if i < nx - 1 else 0.0
                c = row[i]
                hm1 = mw * row[i - 1] + (1.0 - mw) * c if i > 0 else c
                hp1 = me * row[i + 1] + (1.0 - me) * c if i < nx - 1 else c
                s = 0.5 * (hp1 - hm1)
                dmx = max(max(hp1, hm1), c) - c
                dmn = c - min(min(hp1, hm1), c)
                a = min(abs(s), 2.0 * min(dmx, dmn))
                s = a if s > 0.0 else (-a if s < 0.0 else 0.0)
                slp[i] = s if mw * me * m > 0.0 else 0.0
            for i in range(nx):
                mw = (1.0 if active[j, i - 1] else 0.0) if i > 0 else 0.0
                me = (1.0 if active[j, i + 1] else 0.0) if i < nx - 1 else 0.0
                c = row[i]
                hm1 = mw * row[i - 1] + (1.0 - mw) * c if i > 0 else c
                hp1 = me * row[i + 1] + (1.0 - me) * c if i < nx - 1 else c
                sw = slp[i - 1] if i > 0 else 0.0
                se = slp[i + 1] if i < nx - 1 else 0.0
                hw = 0.5 * (hm1 + c) + (sw - slp[i]) / 6.0
                he = 0.5 * (hp1 + c) + (slp[i] - se) / 6.0
                curv = 3.0 * ((hw + he) - 2.0 * c)
                dh = he - hw
                inside = curv > 0.0 and abs(dh) < curv
                if inside and c <= h_min:
                    hw_out[k, j, i] = c
                    he_out[k, j, i] = c
                    continue
                sc = 1.0
                if inside:
                    num = 12.0 * curv * (c - h_min)
                    den = curv * curv + 3.0 * dh * dh
                    if num < den:
                        sc = num / den
                hw_out[k, j, i] = c + sc * (hw - c)
                he_out[k, j, i] = c + sc * (he - c)
    return hw_out, he_out


def _face_flux(k, j, f, u, h, h_w, h_e, dx_t, q_e, q_w, L_f, dt, nx):
    """``_face_flux`` (core/vertical/ale.py:626-648) for one layer of one face."""
    iw = f - 1 if f > 0 else 0
    ie = f if f < nx else nx - 1
    if u >= 0.0:
        c = min(abs(u) * dt / dx_t[j, iw], 1.0)
        hu, ww, ee = h[k, j, iw], h_w[k, j, iw], h_e[k, j, iw]
        curv3, dh, edge, q = (ww + ee) - 2.0 * hu, ww - ee, ee, q_e[k, j, f]
    else:
        c = min(abs(u) * dt / dx_t[j, ie], 1.0)
        hu, ww, ee = h[k, j, ie], h_w[k, j, ie], h_e[k, j, ie]
        curv3, dh, edge, q = (ww + ee) - 2.0 * hu, ee - ww, ww, q_w[k, j, f]
    mean = edge + c * (0.5 * dh + curv3 * (c - 1.5))
    marg = edge + c * (dh + 3.0 * curv3 * (c - 1.0))
    if q < hu:
        mean = min(mean, q)
        marg = min(marg, q)
    return L_f * u * max(mean, 0.0), L_f * max(marg, 0.0)


def ppm_transport_x(h, ut, L, Q_bt, q_e, q_w, active, dt, dx_t, max_iter=20):
    """``_ppm_transport_kernel`` and its wrapper (core/vertical/ale.py:651-764), one face column at a time."""
    act = np.asarray(active, bool)
    hc = np.where(act[None], np.nan_to_num(np.asarray(h, float)), 0.0)
    h_w, h_e = ppm_edges(hc, act)
    h_l = hc.tolist()
    hw_l, he_l = h_w.tolist(), h_e.tolist()
    n, ny, nx = hc.shape
    nf = nx + 1

    class _A:                                   # nested lists indexed like the arrays, floats on access
        def __init__(self, a):
            self.a = a

        def __getitem__(self, idx):
            x = self.a
            for i in idx:
                x = x[i]
            return x

    H, HW, HE = _A(h_l), _A(hw_l), _A(he_l)
    U, QE, QW, DX = _A(np.asarray(ut, float).tolist()), _A(np.asarray(q_e, float).tolist()), \
        _A(np.asarray(q_w, float).tolist()), _A(np.asarray(dx_t, float).tolist())
    Lf, Qb = np.asarray(L, float).tolist(), np.asarray(Q_bt, float).tolist()
    du = [[0.0] * nf for _ in range(ny)]
    lo = [[0.0] * nf for _ in range(ny)]
    hi = [[0.0] * nf for _ in range(ny)]
    err = [[0.0] * nf for _ in range(ny)]
    dsum = [[0.0] * nf for _ in range(ny)]

    def column(j, f, shift):
        iw = f - 1 if f > 0 else 0
        ie = f if f < nx else nx - 1
        s, sd = 0.0, 0.0
        if act[j, iw] and act[j, ie]:
            for k in range(n):
                F, dF = _face_flux(k, j, f, U[k, j, f] + shift if shift is not None else U[k, j, f], H, HW, HE, DX,
                                   QE, QW, Lf[j][f], dt, nx)
                s += F
                sd += dF
        return s, sd

    for j in range(ny):
        for f in range(nf):
            iw = f - 1 if f > 0 else 0
            ie = f if f < nx else nx - 1
            limw, lime = 0.5 * DX[j, iw] / dt, 0.5 * DX[j, ie] / dt
            mnw, mxe = math.inf, -math.inf
            for k in range(n):
                mnw = min(mnw, limw - U[k, j, f])
                mxe = max(mxe, -(lime + U[k, j, f]))
            hi[j][f] = max(min(limw, mnw), 0.0)
            lo[j][f] = min(max(-lime, mxe), 0.0)
            s, sd = column(j, f, None)
            err[j][f] = s - Qb[j][f]
            dsum[j][f] = sd
    iters = 0
    for _ in range(max_iter):
        for j in range(ny):
            for f in range(nf):
                e, d = err[j][f], du[j][f]
                if e > 0.0:
                    hi[j][f] = d
                if e < 0.0:
                    lo[j][f] = d
                step = -e / dsum[j][f] if dsum[j][f] > 0.0 else 0.0
                new = d + step
                if step > 0.0 and new >= hi[j][f]:
                    new = 0.5 * (d + hi[j][f])
                if step < 0.0 and new <= lo[j][f]:
                    new = 0.5 * (d + lo[j][f])
                du[j][f] = new
                s, sd = column(j, f, new)
                err[j][f] = s - Qb[j][f]
                dsum[j][f] = sd
        iters += 1
        done = True
        for j in range(ny):
            for f in range(nf):
                if abs(err[j][f]) > 1e-12 * max(abs(Qb[j][f]), 1.0):
                    done = False
        if done:
            break
    out = np.empty((n, ny, nf))
    for j in range(ny):
        for f in range(nf):
            iw = f - 1 if f > 0 else 0
            ie = f if f < nx else nx - 1
            if not (act[j, iw] and act[j, ie]):
                out[:, j, f] = 0.0
                continue
            Fk = [0.0] * n
            s = 0.0
            for k in range(n):
                Fk[k] = _face_flux(k, j, f, U[k, j, f] + du[j][f], H, HW, HE, DX, QE, QW, Lf[j][f], dt, nx)[0]
                s += Fk[k]
            R = Qb[j][f] - s
            te, tw = 0.0, 0.0
            for k in range(n):
                te += QE[k, j, f]
                tw += QW[k, j, f]
            use_e = R >= 0.0
            tot = te if use_e else tw
            if tot <= 0.0:
                use_e = not use_e
                tot = te if use_e else tw
            den = tot if tot > 0.0 else 1.0
            for k in range(n):
                out[k, j, f] = Fk[k] + R * (QE[k, j, f] if use_e else QW[k, j, f]) / den
    return out, iters
