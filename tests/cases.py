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
