"""
GPU tests of the tracer diffusion kernels: the along-layer Laplacian and the implicit vertical solve are
bit-identical to the oracle's restatement of the reference arithmetic, on a regional grid with land, closed faces,
thin layers and a stack of tracers; the vertical solve also in place, for 1 to 90 levels.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import numpy_ref as R
from pyocean_amd import ops, synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu case]", *a, flush=True)


def _d(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _grid(seed):
    case = synth.regional(nx=130, ny=90, nz=6, nt=3, seed=seed)
    rng = np.random.default_rng(seed)
    act = case.ok.any(axis=0)
    h = np.where(act[None], case.V1 / 1e6, 0.0)
    cx = rng.uniform(0.9, 1.1, (90, 131))
    cy = rng.uniform(0.9, 1.1, (91, 130))
    okx = rng.random((90, 131)) > 0.05
    oky = rng.random((91, 130)) > 0.05
    return case, act, h, cx, cy, okx, oky


@pytest.mark.parametrize("seed", [0, 1])
def test_horizontal_diffusion_matches_the_oracle_bit_for_bit(gpu, seed):
    _say("diffuse_h", seed)
    case, act, h, cx, cy, okx, oky = _grid(seed)
    T = case.C
    Tw = T - 5.0                                             # an anomaly field
    ref = R.diffuse_h_ale(T, Tw, h, act, cx, cy, okx, oky, case.V1, 120.0, case.dt)
    got = ops.diffuse_h(_d(T), _d(Tw), _d(h), _d(act), _d(cx), _d(cy), _d(okx), _d(oky), _d(case.V1), 120.0, case.dt)
    assert np.array_equal(got.cpu().numpy(), ref)
    one = ops.diffuse_h(_d(T[1]), _d(Tw[1]), _d(h), _d(act), _d(cx), _d(cy), _d(okx), _d(oky), _d(case.V1), 120.0,
                        case.dt)
    assert np.array_equal(one.cpu().numpy(), ref[1])


@pytest.mark.parametrize("rates", ["none", "bed", "layers"])
@pytest.mark.parametrize("n,nt", [(17, 3), (1, 2), (16, 5), (50, 10), (64, 4), (65, 3), (90, 2)])
def test_vertical_implicit_matches_the_oracle_bit_for_bit(gpu, rates, n, nt):
    # n <= 64: the register-resident solve (16, 32 and 64 levels); n > 64: the solve through the output array
    _say("vertical_implicit", rates, n, nt)
    rng = np.random.default_rng(4 + n)
    M, dt = 5000, 450.0
    u = rng.normal(5.0, 2.0, (nt, n, M))
    dz = np.maximum(rng.uniform(-1.0, 40.0, (n, M)), 1e-6)
    nu = rng.uniform(1e-6, 5e-2, (n + 1, M))
    r = {"none": None, "bed": rng.uniform(0, 1e-3, M), "layers": rng.uniform(0, 1e-3, (n, M))}[rates]
    tau = rng.normal(0, 1e-4, M) if rates == "bed" else None
    ref = R.vertical_implicit(u, dz, nu, np.zeros(M) if r is None else r, dt, tau)
    got = ops.vertical_implicit(_d(u), _d(dz), _d(nu), None if r is None else _d(r), dt,
                                None if tau is None else _d(tau))
    assert np.array_equal(got.cpu().numpy(), ref)
    inplace = _d(u[0])
    ops.vertical_implicit(inplace, _d(dz), _d(nu), None if r is None else _d(r), dt,
                          None if tau is None else _d(tau), out=inplace)
    assert np.array_equal(inplace.cpu().numpy(), ref[0])


def test_diffusion_at_full_size_conserves_and_keeps_bounds(gpu):
    # the bench workload's size, 1024 x 1024 x 50 with 10 tracers, through size-independent properties: the
    # along-layer fluxes telescope (the content Σ V c is unchanged to round-off) and a uniform field is left exactly
    # as it is; the implicit vertical solve keeps each column's Σ dz c and stays inside the column's range
    _say("diffusion 1024x1024x50 nt=10")
    nt, n, ny, nx = 10, 50, 1024, 1024
    g = torch.Generator(device="cuda").manual_seed(11)
    f64 = dict(dtype=torch.float64, device="cuda")
    act = torch.rand((ny, nx), generator=g, device="cuda") > 0.2
    act8 = act.to(torch.uint8)
    h = torch.where(act[None], 0.01 + 40.0 * torch.rand((n, ny, nx), generator=g, **f64), torch.zeros((), **f64))
    V = h * 1.0e6
    okx = torch.zeros((ny, nx + 1), dtype=torch.uint8, device="cuda")
    okx[:, 1:-1] = act8[:, 1:] & act8[:, :-1]
    oky = torch.zeros((ny + 1, nx), dtype=torch.uint8, device="cuda")
    oky[1:-1] = act8[1:] & act8[:-1]
    cx = 0.9 + 0.2 * torch.rand((ny, nx + 1), generator=g, **f64)
    cy = 0.9 + 0.2 * torch.rand((ny + 1, nx), generator=g, **f64)
    T = 10.0 + torch.randn((nt, n, ny, nx), generator=g, **f64) * act
    dt = 300.0
    out = ops.diffuse_h(T, T, h, act8, cx, cy, okx, oky, V, 100.0, dt)
    c0, c1 = ops.content(T, V), ops.content(out, V)
    scale = ops.content(T.abs(), V)
    assert ((c1 - c0).abs() <= 1e-12 * scale).all(), ((c1 - c0).abs() / scale).max().item()
    uni = torch.full((1, n, ny, nx), 7.25, **f64)
    assert torch.equal(ops.diffuse_h(uni, uni, h, act8, cx, cy, okx, oky, V, 100.0, dt), uni)
    del out, uni
    M = ny * nx
    u = T.view(nt, n, M)
    dz = torch.clamp(h, min=1e-6).view(n, M)
    nu = 1e-4 + 1e-2 * torch.rand((n + 1, M), generator=g, **f64)
    o = ops.vertical_implicit(u, dz, nu, None, dt)
    # wet columns only: a dry column's 1e-6 m layers under the same diffusivity make the tridiagonal system as
    # ill-conditioned as dt nu / dz^2 ~ 1e12 (the reference's arithmetic, reproduced bit for bit, loses 1e-4 there)
    w = act.view(M)
    s0, s1 = (dz[None] * u).sum(1)[:, w], (dz[None] * o).sum(1)[:, w]
    assert ((s1 - s0).abs() <= 1e-12 * (dz[None] * u.abs()).sum(1)[:, w]).all()
    lo, hi = u.min(1, keepdim=True).values, u.max(1, keepdim=True).values
    tol = 1e-12 * ((hi - lo) + hi.abs())
    assert ((o >= lo - tol) & (o <= hi + tol))[:, :, w].all()


@pytest.mark.parametrize("seed", [0, 1])
def test_sigma_diffusion_matches_the_oracle_bit_for_bit(gpu, seed):
    # the sigma branch of diffuse_h: face factors L D / dx with NaN depths read as 0, the layer fractions ds as the
    # face thickness of every face, faces masked where the face depth is below d_min
    _say("diffuse_h sigma", seed)
    rng = np.random.default_rng(20 + seed)
    nt, n, ny, nx = 3, 7, 33, 45
    ds = rng.uniform(0.5, 1.5, n)
    ds /= ds.sum()
    H = rng.uniform(-5.0, 200.0, (ny, nx))
    Du = rng.uniform(-1.0, 150.0, (ny, nx + 1))
    Dv = rng.uniform(-1.0, 150.0, (ny + 1, nx))
    Du[rng.random(Du.shape) < 0.05] = np.nan
    Lu, Lv = rng.uniform(800, 1200, (ny, nx + 1)), rng.uniform(800, 1200, (ny + 1, nx))
    du, dv = rng.uniform(900, 1100, (ny, nx + 1)), rng.uniform(900, 1100, (ny + 1, nx))
    gx, gy = Lu * np.nan_to_num(Du) / du, Lv * np.nan_to_num(Dv) / dv
    okx, oky = np.isfinite(Du) & (Du > 0.1), np.isfinite(Dv) & (Dv > 0.1)
    V = np.where(H > 0, H, 0.0)[None] * ds[:, None, None] * 1e6
    T = rng.normal(10.0, 2.0, (nt, n, ny, nx))
    Tw = T - 8.0
    ref = R.diffuse_h_sigma(T, Tw, ds, gx, gy, okx, oky, V, 50.0, 300.0)
    got = ops.diffuse_h_sigma(_d(T), _d(Tw), ds, _d(gx), _d(gy), _d(okx), _d(oky), _d(V), 50.0, 300.0)
    assert np.array_equal(got.cpu().numpy(), ref)
