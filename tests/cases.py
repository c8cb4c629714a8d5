"""
Small advection inputs shaped like the reference's own test fixtures (no model needed): the z* one-row box of
``tests/test_sigma_ale.py:23-30, 1300-1305`` (20 m deep, dx = dy = 1 km, dt = 30 s) and the open 20 m box of
``:589-604`` (6 layers).  Volumes V = h * dx * dy; transports are built exactly as those tests build them.
"""
from __future__ import annotations

import numpy as np


def step_box(nx=60, nz=4, depth=20.0, dx=1000.0):
    """(h, V) of the one-row z* box: uniform layers depth/nz, shape (nz, 1, nx)."""
    h = np.full((nz, 1, nx), depth / nz)
    return h, h * dx * dx


def open_box(nx=14, ny=12, nz=6, depth=20.0, dx=1000.0):
    h = np.full((nz, ny, nx), depth / nz)
    return h, h * dx * dx


def fused_fct_case(open_=True, seed=7, nx=14, ny=12, nz=6, dt=30.0, thin=True):
    """test_fct_fused_kernels_match_the_numpy_oracle (test_sigma_ale.py:1432-1450): 2-D flow, two tracers, a thin
    (1 mm) and a dry cell, inflow/outflow on all four open edges (or closed)."""
    rng = np.random.default_rng(seed)
    h, _ = open_box(nx, ny, nz)
    h = h.copy()
    if thin:
        h[2, 5, 6] = 1e-3
    V = h * 1e6
    Qx = rng.uniform(-0.4, 0.4, (nz, ny, nx + 1)) * 1e3 * np.concatenate([h[..., :1], h], axis=-1)
    Qy = rng.uniform(-0.4, 0.4, (nz, ny + 1, nx)) * 1e3 * np.concatenate([h[:, :1], h], axis=1)
    if not open_:
        Qx[..., [0, -1]] = 0.0
        Qy[:, [0, -1]] = 0.0
    V1 = V - dt * (Qx[..., 1:] - Qx[..., :-1] + Qy[:, 1:] - Qy[:, :-1])
    ok = V1 > 0
    T = 7.0 + rng.normal(0, 1, (nz, ny, nx))
    C = np.stack([T, T ** 2])
    return dict(C=C, V0=V, V1=V1, Qx=Qx, Qy=Qy, ok=ok, dt=dt, open_mask=15 if open_ else 0)


def fused_rhs_case(seed=7, nx=13, ny=9, nz=6):
    """test_fused_tracer_advection_matches_the_numpy_reference (test_sigma_baroclinic.py:525-544): open W/E/S, a
    closed N wall, dry cells, x, y and z fluxes."""
    rng = np.random.default_rng(seed)
    c = np.stack([rng.normal(10.0, 1.0, (nz, ny, nx)), rng.normal(100.0, 5.0, (nz, ny, nx))])
    ok = rng.random((nz, ny, nx)) > 0.1
    Qx = rng.normal(size=(nz, ny, nx + 1))
    Qy = rng.normal(size=(nz, ny + 1, nx))
    Qy[:, -1, :] = 0.0
    Fz = rng.normal(size=(nz + 1, ny, nx))
    Fz[0] = Fz[-1] = 0.0
    return dict(c=c, Qx=Qx, Qy=Qy, Fz=Fz, ok=ok, open_mask=1 | 2 | 4)


def kat_initial(nx=200):
    """The 1-D known-answer profile (docs/manual/figures/make_tracers.py:45-49): square + Gaussian."""
    x = np.arange(nx) + 0.5
    return np.where((x > 20) & (x < 60), 1.0, 0.0) + np.exp(-((x - 130.0) / 10.0) ** 2)


# ------------------------------------------------------------------------------------------------ ALE layer transport
# Inputs shaped like the reference's tests of core/vertical/ale.py (tests/test_vertical_ale.py:160-165, 361-380,
# 502-516), built with the oracle's transport thickness.
from oracle import numpy_ref as R_  # noqa: E402


def ppm_case(h, u, dt=10.0, dx=100.0, L=10.0, clip=False, H=None, active=None):
    n, ny, nx = h.shape
    active = np.ones((ny, nx), bool) if active is None else active
    H = h.sum(axis=0) if H is None else H
    qe, qw = R_.face_transport_thickness(h, H, active, clip)
    return qe, qw, np.full((ny, nx), dx), np.full((ny, nx + 1), L), active


def ppm_random(seed, n=6, ny=4, nx=9, clip=False, land=False, vanish=False):
    """The reference's random PPM case (tests/test_vertical_ale.py:361-380), built with the oracle."""
    rng = np.random.default_rng(seed)
    H = rng.uniform(50.0, 300.0, (ny, nx)) if clip else np.full((ny, nx), 100.0)
    active = (rng.uniform(size=(ny, nx)) > 0.2) if land else np.ones((ny, nx), bool)
    frac = rng.uniform(0.05, 1.0, (n, ny, nx))
    if vanish:
        frac[rng.uniform(size=frac.shape) < 0.3] = 0.0
        frac[:, 0, 0] = 0.0
        frac[2, 0, 0] = 1.0
    h = frac / frac.sum(0) * H[None]
    ut = rng.normal(0.0, 0.3, (n, ny, nx + 1))
    qe, qw = R_.face_transport_thickness(h, H, active, clip)
    dx_t = rng.uniform(800.0, 1200.0, (ny, nx))
    L = rng.uniform(800.0, 1200.0, (ny, nx + 1))
    F0 = uncorrected(h, ut, L, qe, qw, active, 60.0, dx_t)
    Qbt = F0.sum(0) + rng.normal(0.0, 50.0, (ny, nx + 1))
    return h, ut, L, Qbt, qe, qw, active, dx_t


def uncorrected(h, u, L, qe, qw, active, dt, dx_t):
    """Layer fluxes without the Newton correction (zero on closed faces)."""
    hc = np.where(active[None], h, 0.0)
    hw, he = R_.ppm_edges(hc, active)
    F, _ = R_._ppm_face_flux(u, hc, hw, he, dx_t, qe, qw, L, dt)
    nx = h.shape[-1]
    f = np.arange(nx + 1)
    open_ = active[:, np.maximum(f - 1, 0)] & active[:, np.minimum(f, nx - 1)]
    return np.where(open_[None], F, 0.0)


def thin_divergent_cell():
    dx, dt = 5000.0, 300.0
    h = np.array([[[50.0, 50.0, 50.0]], [[0.08, 0.001, 49.9]], [[50.0, 50.0, 50.0]]])
    area = np.full((1, 3), dx * dx)
    Qx = np.zeros((3, 1, 4))
    Qx[1, 0, 1] = -0.459 * 0.001 * dx * dx / dt
    Qx[1, 0, 2] = +0.553 * 0.001 * dx * dx / dt
    Qx[0, 0, 1:3] = [0.2, 0.3]
    Qx[2, 0, 1:3] = [-0.4, 0.1]
    return h, Qx, np.zeros((3, 2, 3)), area, dt, np.ones((1, 3), bool)


def limiter_case(seed, n=5, ny=9, nx=11, periodic=(False, False)):
    """Thin layers under strong, partly counter-flowing transports in x and y, some land."""
    rng = np.random.default_rng(seed)
    act = rng.uniform(size=(ny, nx)) > 0.1
    h = rng.uniform(1.0, 20.0, (n, ny, nx))
    h[rng.uniform(size=h.shape) < 0.25] = rng.uniform(1e-4, 5e-3)
    h = np.where(act[None], h, 0.0)
    area = rng.uniform(0.8e6, 1.2e6, (ny, nx))
    dt = 60.0
    Qx = rng.normal(0.0, 1.0, (n, ny, nx + 1)) * 0.25 * h.mean() * 1000.0 * 0.5
    Qy = rng.normal(0.0, 1.0, (n, ny + 1, nx)) * 0.25 * h.mean() * 1000.0 * 0.5
    f = np.arange(nx + 1)
    openx = act[:, np.maximum(f - 1, 0)] & act[:, np.minimum(f, nx - 1)]
    g = np.arange(ny + 1)
    openy = act[np.maximum(g - 1, 0)] & act[np.minimum(g, ny - 1)]
    Qx = np.where(openx[None], Qx, 0.0)
    Qy = np.where(openy[None], Qy, 0.0)
    if periodic[0]:
        Qx[..., -1] = Qx[..., 0]
    if periodic[1]:
        Qy[:, -1] = Qy[:, 0]
    return h, Qx, Qy, area, dt, act


# ------------------------------------------------------------------------------------------------ vertical remap
def remap_case(seed, n0=50, n1=50, ny=9, nx=23, nf=3):
    """Columns of every kind the remap meets: thick, z*-like with vanished bed layers, interleaved vanished layers
    (0, 1e-10, 1e-3 m), 1-3 active cells (PCM / PLM), all vanished (dry), a dry mask, target totals off by +-1e-9 m
    and +-1e-13 m, vanished targets.  Returns h_src (n0, ny, nx), f (nf, n0, ny, nx), h_tgt (n1, ny, nx), wet."""
    rng = np.random.default_rng(seed)
    M = ny * nx
    hs = np.empty((n0, M))
    for c in range(M):
        kind = c % 8
        if kind == 0:
            col = rng.uniform(0.5, 40.0, n0)
        elif kind == 1:
            nv = int(rng.integers(1, max(n0 - 4, 2)))
            col = np.r_[np.zeros(nv), rng.uniform(2.0, 30.0, n0 - nv)]
        elif kind == 2:
            col = np.where(rng.random(n0) < 0.7, rng.choice([0.0, 1e-10, 1e-3], n0), rng.uniform(0.5, 50.0, n0))
        elif kind in (3, 4, 5):
            col = np.zeros(n0)
            na = min(kind - 2, n0)
            col[rng.choice(n0, na, replace=False)] = rng.uniform(0.5, 10.0, na)
        elif kind == 6:
            col = np.zeros(n0) if c % 16 == 6 else rng.uniform(0.0, 3.0, n0)
        else:
            col = np.where(rng.random(n0) < 0.2, 0.0, rng.uniform(0.1, 20.0, n0))
        hs[:, c] = col
    ht = rng.uniform(0.1, 2.0, (n1, M))
    ht[rng.random((n1, M)) < 0.1] = 0.0                                  # vanished targets
    tot_s, tot_t = hs.sum(0), ht.sum(0)
    ht = np.where(tot_t > 0, ht * (tot_s / np.where(tot_t > 0, tot_t, 1.0)), ht)
    ht[-1, 1::5] += 1e-9
    ht[-1, 2::5] += 1e-13
    ht[-1, 3::7] -= np.minimum(1e-9, ht[-1, 3::7])
    f = np.cumsum(rng.normal(0.0, 1.0, (nf, n0, M)), axis=1) + rng.uniform(5.0, 20.0, (nf, 1, M))
    f[:, :, 7::11] = 12.5                                                 # uniform columns
    wet = rng.random(M) > 0.1
    return hs.reshape(n0, ny, nx), f.reshape(nf, n0, ny, nx), ht.reshape(n1, ny, nx), wet.reshape(ny, nx)
