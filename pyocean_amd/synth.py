"""
Seeded synthetic advection workloads (inputs of one ``advect_fct`` / ``advect`` call), sized like the benchmark
cases.  All arrays follow the reference layout: stack C (nt, N, ny, nx), cells V0, V1, ok (N, ny, nx), faces
Qx (N, ny, nx+1), Qy (N, ny+1, nx), Fz (N+1, ny, nx), k = 0 at the bed, FP64.

* ``solid_body`` - 128 x 128 x 20 Cartesian box, dx = dy = 1 km, flat 200 m, 10 m z* layers, a rigid rotation
  inside the inscribed disk (stream function psi = Omega/2 min(r, R)^2, constant on every edge, so the layer
  transports are exactly non-divergent and the edges closed), Courant 0.4; a Gaussian blob (sigma 8 cells) and a
  slotted cylinder (0/1).
* ``regional`` - curvilinear grid with +-5 % smooth metric noise, ~20 % land from Gaussian-filtered bathymetry
  (10-2000 m), z* layers following eta (+-0.5 m), horizontal velocity from a random stream function with a k^-3
  spectrum (rms 0.3 m/s), face transports with the thinner adjacent layer, the Lagrangian volume
  V1 = V0 - dt div Q (Fz = 0, the ALE way), dt from Courant 0.3; tracers [T, S, 8 passive].
* ``regional_sigma`` - the same grid with V1 = V0 and the continuity residual carried by Fz (the z-sweep case).

The layer rule is z* (``zstar_column``, core/vertical/regrid.py:43-59); the transports follow the layer-flux
pattern of core/components/layer_transport.py:111-156 (Lagrangian volumes, ok = active & V1 > 0).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

H_MIN = 1e-3            # minimum z* layer thickness [m] (core/vertical/regrid.py:40)


@dataclass
class Case:
    """One advection call's inputs."""
    C: np.ndarray
    V0: np.ndarray
    V1: np.ndarray
    Qx: np.ndarray
    Qy: np.ndarray
    ok: np.ndarray
    dt: float
    Fz: np.ndarray | None = None
    inflow: np.ndarray | None = None
    open_edges: tuple = ()
    info: dict = field(default_factory=dict)

    @property
    def shape(self):
        return self.C.shape

    def cells(self) -> int:
        return int(np.prod(self.C.shape))


# ------------------------------------------------------------------------------------------------ helpers
def _smooth_noise(rng, ny, nx, sigma):
    """Gaussian-filtered white noise (periodic FFT filter) normalised to [-1, 1]."""
    w = rng.standard_normal((ny, nx))
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.rfftfreq(nx)[None, :]
    g = np.exp(-2.0 * (np.pi * sigma) ** 2 * (kx ** 2 + ky ** 2))
    f = np.fft.irfft2(np.fft.rfft2(w) * g, s=(ny, nx))
    f -= f.min()
    f /= max(f.max(), 1e-300)
    return 2.0 * f - 1.0


def _spectral_streamfunction(rng, ny, nx, slope=-3.0):
    """Random-phase field on (ny, nx) corners with energy spectrum ~ k^slope (unnormalised)."""
    ky = np.fft.fftfreq(ny)[:, None] * ny
    kx = np.fft.rfftfreq(nx)[None, :] * nx
    kk = np.sqrt(kx ** 2 + ky ** 2)
    kk[0, 0] = 1.0
    # velocity energy E(k) ~ k^slope -> |psi_hat| ~ k^((slope - 1)/2 - 1)
    amp = kk ** ((slope - 1.0) / 2.0 - 1.0)
    amp[0, 0] = 0.0
    phase = rng.uniform(0.0, 2.0 * np.pi, amp.shape)
    return np.fft.irfft2(amp * np.exp(1j * phase), s=(ny, nx))


def zstar_layers(H, eta, dz_top, h_min=H_MIN):
    """
    z* layer thicknesses (N, ny, nx), bed first, for depth H > 0 and surface eta, nominal thicknesses dz_top
    (surface first): interfaces stretched by s = (H + eta)/H, bottom interface at -H, thin layers raised to h_min
    (core/vertical/regrid.py:43-59, vectorised over columns).
    """
    n = dz_top.size
    D = H + eta
    hm = np.minimum(h_min, D / n)
    s = D / H
    e = np.empty((n + 1,) + H.shape)
    e[0] = eta
    for k in range(n):
        e[k + 1] = e[k] - s * dz_top[k]
    e[n] = -H
    for k in range(n - 1, -1, -1):
        e[k] = np.where(e[k] < e[k + 1] + hm, e[k + 1] + hm, e[k])
    zw = e[::-1]                                   # bed first
    return zw[1:] - zw[:-1], zw


# ------------------------------------------------------------------------------------------------ solid body
def solid_body(n=128, nz=20, nt=2, depth=200.0, dx=1000.0, courant=0.4, dt=100.0, seed=0) -> Case:
    """Rigid rotation of a Gaussian blob and a slotted cylinder (the CPU-runnable case)."""
    ny = nx = n
    h = np.full((nz, ny, nx), depth / nz)
    area = dx * dx
    V = h * area
    xc = yc = 0.5 * n * dx
    R = 0.5 * n * dx
    xq = np.arange(nx + 1) * dx
    yq = np.arange(ny + 1) * dx
    r2 = (xq[None, :] - xc) ** 2 + (yq[:, None] - yc) ** 2
    psi = 0.5 * np.minimum(r2, R * R)              # per unit Omega; constant (R^2/2) on every edge
    ux = -(psi[1:, :] - psi[:-1, :]) / dx          # (ny, nx+1) u per Omega
    vy = (psi[:, 1:] - psi[:, :-1]) / dx           # (ny+1, nx)
    umax = max(np.abs(ux).max(), np.abs(vy).max())
    omega = courant * dx / (dt * umax)
    Qx = np.broadcast_to(omega * ux * dx, (nz, ny, nx + 1)) * (depth / nz)
    Qy = np.broadcast_to(omega * vy * dx, (nz, ny + 1, nx)) * (depth / nz)
    Qx, Qy = np.ascontiguousarray(Qx), np.ascontiguousarray(Qy)
    div = Qx[..., 1:] - Qx[..., :-1] + Qy[:, 1:, :] - Qy[:, :-1, :]
    V1 = V - dt * div
    ok = V1 > 0
    ic = np.arange(nx) + 0.5
    jc = np.arange(ny) + 0.5
    X, Y = np.meshgrid(ic, jc)
    blob = np.exp(-0.5 * ((X - 0.5 * n) ** 2 + (Y - 0.25 * n) ** 2) / 8.0 ** 2)
    rr = np.sqrt((X - 0.5 * n) ** 2 + (Y - 0.75 * n) ** 2)
    cyl = ((rr <= n / 8) & ~((np.abs(X - 0.5 * n) <= 2) & (Y <= 0.75 * n + 8))).astype(float)
    C = np.stack([np.broadcast_to(blob, (nz, ny, nx)), np.broadcast_to(cyl, (nz, ny, nx))])[:nt].copy()
    return Case(C=C, V0=V, V1=V1, Qx=Qx, Qy=Qy, ok=ok, dt=float(dt),
                info=dict(name=f"solid_body {nx}x{ny}x{nz} nt={nt}", omega=omega, steps=100))


# ------------------------------------------------------------------------------------------------ regional
def _regional_grid(nx, ny, nz, seed, land_frac=0.2, hmax=2000.0, hmin=10.0):
    rng1 = np.random.default_rng(seed + 1)
    sig = max(2.0, 32.0 * min(nx, ny) / 1024.0)
    m = _smooth_noise(rng1, ny, nx, sig)
    dx_t = 1000.0 * (1.0 + 0.05 * m)
    dy_t = 1000.0 * (1.0 - 0.05 * m[::-1, :])
    area = dx_t * dy_t
    rng2 = np.random.default_rng(seed + 2)
    b = _smooth_noise(rng2, ny, nx, max(2.0, 48.0 * min(nx, ny) / 1024.0))
    thr = np.quantile(b, land_frac)
    # linear map with H = 0 at the land quantile, H = hmax at the maximum
    H = hmax * (b - thr) / (b.max() - thr)
    active = H > 0.0
    H = np.where(active, np.maximum(H, hmin), 0.0)
    rng3 = np.random.default_rng(seed + 3)
    eta = 0.5 * _smooth_noise(rng3, ny, nx, max(2.0, 24.0 * min(nx, ny) / 1024.0)) * active
    dz_top = np.full(nz, hmax / nz)
    Hs = np.where(active, H, 1.0)
    h, zw = zstar_layers(Hs, np.where(active, eta, 0.0), dz_top)
    h = np.where(active[None], h, 0.0)
    zc = 0.5 * (zw[1:] + zw[:-1])                  # layer centres, negative down
    return dict(dx_t=dx_t, dy_t=dy_t, area=area, H=H, active=active, eta=eta, h=h, zc=zc)


def _regional_transports(g, seed, urms=0.3):
    h, active = g["h"], g["active"]
    nz, ny, nx = h.shape
    rng4 = np.random.default_rng(seed + 4)
    psi = _spectral_streamfunction(rng4, ny + 1, nx + 1)
    # face lengths: average of the adjacent cell metrics (edges: the cell's own)
    dy_t, dx_t = g["dy_t"], g["dx_t"]
    dyu = np.concatenate([dy_t[:, :1], 0.5 * (dy_t[:, 1:] + dy_t[:, :-1]), dy_t[:, -1:]], axis=1)
    dxv = np.concatenate([dx_t[:1], 0.5 * (dx_t[1:] + dx_t[:-1]), dx_t[-1:]], axis=0)
    u = -(psi[1:, :] - psi[:-1, :]) / dyu          # (ny, nx+1)
    v = (psi[:, 1:] - psi[:, :-1]) / dxv           # (ny+1, nx)
    wet_u = np.zeros((ny, nx + 1), bool)
    wet_u[:, 1:-1] = active[:, 1:] & active[:, :-1]
    wet_v = np.zeros((ny + 1, nx), bool)
    wet_v[1:-1] = active[1:] & active[:-1]
    u *= wet_u
    v *= wet_v
    rms = np.sqrt((np.sum(u ** 2) + np.sum(v ** 2)) / max(1, wet_u.sum() + wet_v.sum()))
    u *= urms / rms
    v *= urms / rms
    hu = np.zeros((nz, ny, nx + 1))
    hu[..., 1:-1] = np.minimum(h[..., 1:], h[..., :-1])
    hv = np.zeros((nz, ny + 1, nx))
    hv[:, 1:-1] = np.minimum(h[:, 1:], h[:, :-1])
    Qx = u[None] * dyu[None] * hu
    Qy = v[None] * dxv[None] * hv
    return Qx, Qy, u, v, dyu, dxv


def _regional_tracers(g, nt, seed, active):
    nz, ny, nx = g["h"].shape
    zc = g["zc"]
    depth = np.clip(-zc, 0.0, 2000.0)
    C = np.empty((nt, nz, ny, nx))
    if nt >= 1:
        rng5 = np.random.default_rng(seed + 5)
        C[0] = 28.0 - 26.0 * depth / 2000.0 + rng5.uniform(-0.5, 0.5, (nz, ny, nx))
    if nt >= 2:
        C[1] = 34.0 + 2.0 * depth / 2000.0
    for m in range(2, nt):
        C[m] = np.random.default_rng(seed + 4 + m).random((nz, ny, nx))
    return C


def regional(nx=1024, ny=1024, nz=50, nt=10, courant=0.3, seed=0, sigma=False) -> Case:
    """
    The regional case.  ``sigma``: V1 = V0 with the continuity residual in Fz (cumulative from the bed,
    core/components/layer_transport.py:85-88) instead of the Lagrangian V1.
    """
    g = _regional_grid(nx, ny, nz, seed)
    Qx, Qy, u, v, dyu, dxv = _regional_transports(g, seed)
    dx_t, dy_t, area, h, active = g["dx_t"], g["dy_t"], g["area"], g["h"], g["active"]
    # time step from the face Courant number
    cx = np.abs(u[:, 1:-1]) / (0.5 * (dx_t[:, 1:] + dx_t[:, :-1]))
    cy = np.abs(v[1:-1]) / (0.5 * (dy_t[1:] + dy_t[:-1]))
    dt = courant / max(cx.max(), cy.max())
    V0 = h * area[None]
    div = Qx[..., 1:] - Qx[..., :-1] + Qy[:, 1:, :] - Qy[:, :-1, :]
    out_frac = np.where(V0 > 0, dt * (np.maximum(Qx[..., 1:], 0) - np.minimum(Qx[..., :-1], 0)
                                      + np.maximum(Qy[:, 1:], 0) - np.minimum(Qy[:, :-1], 0))
                        / np.where(V0 > 0, V0, 1.0), 0.0)
    if out_frac.max() >= 1.0:                      # limit_outflow's condition (core/vertical/ale.py:331)
        dt *= 0.9 / out_frac.max()
        out_frac *= 0.9 / out_frac.max()
    Fz = None
    if sigma:
        # the column's net convergence changes every layer in proportion to its thickness (a sigma-like surface
        # motion); the rest of each layer's divergence leaves through the interfaces, cumulated from the bed
        D = np.maximum(h.sum(axis=0), 1e-300)
        share = div.sum(axis=0)[None] * (h / D[None])
        Fz = np.zeros((nz + 1, ny, nx))
        Fz[1:] = np.cumsum(share - div, axis=0)
        Fz[nz] = 0.0
        # the interfaces carry the divergence of every layer below them, so a thin layer can see far more vertical
        # than horizontal flux: the time step also keeps each cell's outflow through all six faces below 0.6 of
        # its volume (the horizontal rule's bound), which keeps the RK3 step stable
        out3 = np.where(V0 > 0, dt * (np.maximum(Qx[..., 1:], 0) - np.minimum(Qx[..., :-1], 0)
                                      + np.maximum(Qy[:, 1:], 0) - np.minimum(Qy[:, :-1], 0)
                                      + np.maximum(Fz[1:], 0) - np.minimum(Fz[:-1], 0))
                        / np.where(V0 > 0, V0, 1.0), 0.0)
        if out3.max() > 0.6:
            dt *= 0.6 / out3.max()
        V1 = V0 - dt * share
    else:
        V1 = V0 - dt * div
    ok = active[None] & (V1 > 0)
    C = _regional_tracers(g, nt, seed, active)
    info = dict(name=f"regional {nx}x{ny}x{nz} nt={nt}", land=float(1.0 - active.mean()),
                wet_cells=int(ok.sum()), max_outflow_fraction=float(out_frac.max()), dt=float(dt), steps=200)
    return Case(C=C, V0=V0, V1=V1, Qx=Qx, Qy=Qy, ok=ok, dt=float(dt), Fz=Fz, info=info)


def batch_shapes(count=64, seed=7):
    """Domain shapes of the batched case: nx, ny log-uniform in [96, 768], nz in {20, 30, 50, 75}."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nx = int(round(np.exp(rng.uniform(np.log(96), np.log(768)))))
        ny = int(round(np.exp(rng.uniform(np.log(96), np.log(768)))))
        nz = int(rng.choice([20, 30, 50, 75]))
        out.append((nx, ny, nz))
    return out


# ------------------------------------------------------------------------------------------------ ALE producer inputs
@dataclass
class AleInputs:
    """Inputs of one ALE layer-transport step (core/components/layer_transport.py:111-156) on the regional grid."""
    h: np.ndarray             # (N, ny, nx) layer thickness at step n, bed first
    H: np.ndarray             # (ny, nx) rest depth
    active: np.ndarray        # (ny, nx)
    area: np.ndarray
    dx_t: np.ndarray
    dy_t: np.ndarray
    Lu: np.ndarray            # (ny, nx+1) u-face lengths
    Lv: np.ndarray            # (ny+1, nx)
    u_new: np.ndarray         # (N, ny, nx+1)
    u_old: np.ndarray
    v_new: np.ndarray         # (N, ny+1, nx)
    v_old: np.ndarray
    wet_u: np.ndarray         # (ny, nx+1) faces wet in the last free-surface step
    wet_v: np.ndarray
    wu: np.ndarray            # (N, ny, nx+1) face fractions
    wv: np.ndarray
    Du: np.ndarray            # (ny, nx+1) face depths
    Dv: np.ndarray
    dt: float
    theta: float = 0.6
    theta_bc: float = 0.5
    info: dict = field(default_factory=dict)


def _face_fractions(hf):
    """w_k = h_k,f / Σ_k h_k,f, uniform 1/N where the face holds no water (core/vertical/ale.py:256-261)."""
    tot = hf.sum(axis=0)
    wet = tot > 0
    return np.where(wet[None], hf / np.where(wet, tot, 1.0)[None], 1.0 / hf.shape[0])


def ale_inputs(nx=1024, ny=1024, nz=50, courant=0.3, shear=0.1, seed=0) -> AleInputs:
    """
    The regional grid's z* layers and a velocity field for the ALE layer transport: the barotropic stream-function
    velocity of ``regional`` plus a baroclinic shear (``shear`` x the rms, a smooth random field times a profile
    decaying with depth) and, for the old step, the same field turned by a few per cent; face fractions and depths
    from the mean face thickness (q_E + q_W)/2 of the z* layers (clipped at the shallower bed).  dt from the face
    Courant number ``courant`` of the new velocity.
    """
    g = _regional_grid(nx, ny, nz, seed)
    _, _, u, v, dyu, dxv = _regional_transports(g, seed)
    h, active, H = g["h"], g["active"], g["H"]
    rng = np.random.default_rng(seed + 11)
    prof = np.exp(g["zc"] / 500.0)                                     # (nz, ny, nx), 1 at the surface
    pu = np.concatenate([prof[..., :1], prof], axis=-1)
    pv = np.concatenate([prof[:, :1], prof], axis=1)
    wet_u = np.zeros((ny, nx + 1), bool)
    wet_u[:, 1:-1] = active[:, 1:] & active[:, :-1]
    wet_v = np.zeros((ny + 1, nx), bool)
    wet_v[1:-1] = active[1:] & active[:-1]
    urms = np.sqrt(np.mean(u[wet_u] ** 2))
    su = _smooth_noise(rng, ny, nx + 1, max(2.0, 16.0 * min(nx, ny) / 1024.0))
    sv = _smooth_noise(rng, ny + 1, nx, max(2.0, 16.0 * min(nx, ny) / 1024.0))
    u_new = (u[None] + shear * urms * su[None] * (pu - pu.mean(axis=0))) * wet_u[None]
    v_new = (v[None] + shear * urms * sv[None] * (pv - pv.mean(axis=0))) * wet_v[None]
    turn = 0.03
    u_old = (1.0 - turn) * u_new + turn * np.roll(u_new, 1, axis=-1) * wet_u[None]
    v_old = (1.0 - turn) * v_new + turn * np.roll(v_new, 1, axis=-2) * wet_v[None]
    # mean face thickness of the z* layers clipped at the shallower bed (face_transport_thickness, clip=True)
    hz = np.where(active[None], h, 0.0)

    def mean_face(hc, Hc, act):
        n_, a_, b_ = hc.shape
        f = np.arange(b_ + 1)
        iw, ie = np.maximum(f - 1, 0), np.minimum(f, b_ - 1)
        op = act[:, iw] & act[:, ie]
        z_bed = np.maximum(-Hc[:, iw], -Hc[:, ie])
        qe = np.minimum(hc[..., iw], np.maximum(0.0, (-Hc[:, iw] + np.cumsum(hc[..., iw], axis=0)) - z_bed))
        qw = np.minimum(hc[..., ie], np.maximum(0.0, (-Hc[:, ie] + np.cumsum(hc[..., ie], axis=0)) - z_bed))
        return np.where(op[None], 0.5 * (qe + qw), 0.0)

    hfu = mean_face(hz, H, active)
    hfv = mean_face(hz.swapaxes(1, 2), H.T, active.T).swapaxes(1, 2)
    dx_t, dy_t = g["dx_t"], g["dy_t"]
    cx = np.abs(u_new[:, :, 1:-1]).max(axis=0) / (0.5 * (dx_t[:, 1:] + dx_t[:, :-1]))
    cy = np.abs(v_new[:, 1:-1]).max(axis=0) / (0.5 * (dy_t[1:] + dy_t[:-1]))
    dt = courant / max(cx.max(), cy.max())
    info = dict(name=f"ale regional {nx}x{ny}x{nz}", land=float(1.0 - active.mean()), dt=float(dt))
    return AleInputs(h=h, H=H, active=active, area=g["area"], dx_t=dx_t, dy_t=dy_t, Lu=dyu, Lv=dxv,
                     u_new=u_new, u_old=u_old, v_new=v_new, v_old=v_old, wet_u=wet_u, wet_v=wet_v,
                     wu=_face_fractions(hfu), wv=_face_fractions(hfv), Du=hfu.sum(axis=0), Dv=hfv.sum(axis=0),
                     dt=float(dt), info=info)
