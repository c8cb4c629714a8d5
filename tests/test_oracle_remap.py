"""
Pins of the oracle's vertical regrid (z*, sigma) and conservative remap (PCM, PLM, PPM_H4) (CPU).

* The reference's known answers (tests/test_vertical_remap.py and tests/test_vertical_regrid.py of the reference):
  identity, conservation, the PCM values, vanished targets, round-off thickness mismatches, linear exactness of
  PLM and quadratic exactness of PPM_H4, no new extrema, order of accuracy between the schemes, short columns,
  z*-like and interleaved vanished layers, the dry column, and MOM6's closed-form end values
  ([10, 12.642, 13.358, 14.338, 15.052, 15.61, 16], from MOM6 end_value_h4); the z* and sigma columns.
* Mathematics: every target mean is the integral of the source cells' reconstruction over the target cell
  (quadrature of the piecewise polynomial), and the target content equals the source content.
"""
from __future__ import annotations

import numpy as np
import pytest
from scipy import integrate

from oracle import cref

PCM, PLM, PPM = cref.PCM, cref.PLM, cref.PPM_H4
DELR = np.array([1.0, 2.0, 3.0, 5.0, 8.0, 12.0, 20.0, 30.0, 50.0, 70.0])   # surface first, sum 201
H_MIN = 1e-3


def _remap(h0, u0, h1, scheme):
    return cref.remap_column(h0, u0, h1, scheme)


@pytest.mark.parametrize("scheme", [PCM, PLM, PPM])
def test_identity_when_grid_unchanged(scheme):
    rng = np.random.default_rng(0)
    h = rng.uniform(0.5, 3.0, 12)
    u = rng.normal(10, 2, 12)
    np.testing.assert_allclose(_remap(h, u, h, scheme), u, rtol=1e-14)


@pytest.mark.parametrize("scheme", [PCM, PLM, PPM])
def test_conservation_to_roundoff(scheme):
    rng = np.random.default_rng(1)
    h0 = rng.uniform(0.0, 4.0, 20)
    h0[[3, 4, 11]] = 0.0
    u0 = rng.normal(15, 3, 20)
    h1 = rng.uniform(0.1, 2.0, 15)
    h1 *= h0.sum() / h1.sum()
    u1 = _remap(h0, u0, h1, scheme)
    assert abs((h1 * u1).sum() - (h0 * u0).sum()) <= 1e-12 * abs((h0 * u0).sum())


def test_pcm_values_and_vanished_target():
    np.testing.assert_allclose(_remap([1, 1], [1.0, 3.0], [2.0], PCM), [2.0])
    np.testing.assert_allclose(_remap([1, 1], [1.0, 3.0], [0.5, 1.5], PCM), [1.0, (0.5 + 3.0) / 1.5])
    u1 = _remap([1.0, 1.0], [1.0, 3.0], [1.0, 0.0, 1.0], PCM)
    assert np.isfinite(u1).all()
    np.testing.assert_allclose(u1[[0, 2]], [1.0, 3.0])
    assert 1.0 <= u1[1] <= 3.0


def test_thickness_mismatches_keep_the_content():
    h0, u0 = np.array([1.0, 2.0, 3.0]), np.array([4.0, 5.0, 6.0])
    u1 = _remap(h0, u0, np.array([3.0, 3.0 + 1e-13]), PCM)
    assert np.all((u1 >= 4.0 - 1e-12) & (u1 <= 6.0 + 1e-12))
    assert abs((np.array([3.0, 3.0 + 1e-13]) * u1).sum() - (h0 * u0).sum()) < 1e-11
    h0b = np.array([1.0, 2.0, 3.0 + 1e-9])
    u1 = _remap(h0b, u0, np.array([3.0, 3.0]), PCM)
    assert abs((np.array([3.0, 3.0]) * u1).sum() - (h0b * u0).sum()) < 1e-12
    for scheme in (PCM, PPM):
        h1 = np.array([3.0, 3.0 + 1e-9])
        u1 = _remap(h0, u0, h1, scheme)
        assert abs((h1 * u1).sum() - (h0 * u0).sum()) < 1e-12


def test_plm_exact_for_linear_profile_interior():
    h = np.array([1.0, 0.5, 2.0, 1.0, 1.5])
    z = np.concatenate([[0.0], np.cumsum(h)])
    u = 2 + 3 * 0.5 * (z[1:] + z[:-1])
    coef = cref.reconstruct(h, u, PLM)
    for k in (1, 2, 3):
        np.testing.assert_allclose(coef[k, 0], 2 + 3 * z[k], rtol=1e-12)
        np.testing.assert_allclose(coef[k, 0] + coef[k, 1], 2 + 3 * z[k + 1], rtol=1e-12)
    np.testing.assert_allclose(coef[[0, 4], 1], 0.0)


def test_ppm_exact_for_quadratic_interior():
    h = np.array([1.0, 0.7, 1.3, 0.9, 1.1, 0.8, 1.2])
    z = np.concatenate([[0.0], np.cumsum(h)]) + 1.0
    F = lambda x: x + x ** 2 / 2 + x ** 3 / 3                     # noqa: E731
    u = (F(z[1:]) - F(z[:-1])) / h
    coef = cref.reconstruct(h, u, PPM)
    for k in range(2, 5):
        np.testing.assert_allclose(coef[k, 0], 1 + z[k] + z[k] ** 2, rtol=1e-10)
        np.testing.assert_allclose(coef[k].sum(), 1 + z[k + 1] + z[k + 1] ** 2, rtol=1e-10)


@pytest.mark.parametrize("scheme", [PCM, PLM, PPM])
def test_no_new_extrema(scheme):
    rng = np.random.default_rng(3 if scheme != PPM else 4)
    n0 = 30 if scheme != PPM else 40
    h0 = rng.uniform(0.2, 2.0, n0)
    u0 = np.where(np.arange(n0) < n0 // 2, 10.0, 20.0) + rng.normal(0, 0.5 if scheme != PPM else 0.3, n0)
    h1 = rng.uniform(0.1, 1.0, 50 if scheme != PPM else 60)
    h1 *= h0.sum() / h1.sum()
    u1 = _remap(h0, u0, h1, scheme)
    assert u1.min() >= u0.min() - 1e-12 and u1.max() <= u0.max() + 1e-12


def test_higher_order_schemes_diffuse_less():
    h0 = np.full(40, 1.0)
    u0 = np.sin(2 * np.pi * (np.arange(40) + 0.5) / 40)
    hs = np.concatenate([[0.3], np.full(39, 1.0), [0.7]])
    err = {}
    for scheme in (PCM, PLM, PPM):
        u = u0.copy()
        for _ in range(10):
            u = _remap(hs, _remap(h0, u, hs, scheme), h0, scheme)
        err[scheme] = np.abs(u - u0).max()
    assert err[PLM] < 0.5 * err[PCM] and err[PPM] < err[PLM]


def test_short_columns_reduce_order_and_dry_columns_keep_their_values():
    for n in (1, 2, 3):
        h0 = np.ones(n)
        np.testing.assert_allclose(_remap(h0, np.arange(n, dtype=float), h0, PPM), np.arange(n), rtol=1e-14)
    for scheme in (PCM, PLM, PPM):
        u0 = np.array([20.0, 21.0, 22.0, 23.0])
        np.testing.assert_array_equal(_remap(np.zeros(4), u0, np.zeros(4), scheme), u0)
        np.testing.assert_allclose(_remap(np.zeros(4), u0, np.zeros(2), scheme), 21.5)


def test_vanished_layers_in_the_end_stencil_give_mom6s_end_values():
    u1 = _remap([0, 1, 0, 0, 2, 3, 1], [5, 10, 11, 12, 13, 15, 16], np.ones(7), PPM)
    np.testing.assert_allclose(u1, [10, 12.642, 13.358, 14.338, 15.052, 15.61, 16], atol=1e-3)


def test_zstar_like_and_interleaved_vanished_layers():
    h0 = np.r_[np.full(44, 0.0), np.linspace(5, 20, 6)]
    u0 = np.r_[np.full(44, 12.0), np.linspace(12, 18, 6)]
    h1 = np.r_[np.full(44, 1e-3), np.linspace(5, 20, 6)]
    h1[-1] -= 44e-3
    for scheme in (PCM, PLM, PPM):
        u1 = _remap(h0, u0, h1, scheme)
        assert np.isfinite(u1).all()
        assert abs((h1 * u1).sum() - (h0 * u0).sum()) <= 1e-12 * (h0 * u0).sum()
    rng = np.random.default_rng(5)
    for _ in range(200):
        n = 50
        h0 = np.where(rng.random(n) < 0.85, rng.choice([0.0, 1e-10, 1e-3], n), rng.uniform(0.5, 50.0, n))
        u0 = rng.normal(15, 3, n)
        h1 = rng.uniform(0.1, 2.0, n)
        h1 *= h0.sum() / h1.sum()
        u1 = _remap(h0, u0, h1, PPM)
        assert np.isfinite(u1).all()
        assert abs((h1 * u1).sum() - (h0 * u0).sum()) <= 1e-11 * abs((h0 * u0).sum())


def test_remap_columns_matches_single_columns_and_copies_dry_ones():
    rng = np.random.default_rng(5)
    N, ny, nx = 8, 3, 4
    h0 = rng.uniform(0.5, 2.0, (N, ny, nx))
    f = rng.normal(10, 1, (N, ny, nx))
    h1 = rng.uniform(0.5, 2.0, (N, ny, nx))
    h1 *= h0.sum(0) / h1.sum(0)
    wet = np.ones((ny, nx), bool)
    wet[0, 0] = False
    out = cref.remap_columns(h0, f, h1, PPM, wet)
    np.testing.assert_array_equal(out[:, 1, 2], _remap(h0[:, 1, 2], f[:, 1, 2], h1[:, 1, 2], PPM))
    np.testing.assert_array_equal(out[:, 0, 0], f[:, 0, 0])
    stack = cref.remap_columns(h0, np.stack([f, 2 * f]), h1, PPM, wet)
    np.testing.assert_array_equal(stack[0], out)
    h1b = h1[:5] * (h1.sum(0) / h1[:5].sum(0))[None]
    assert np.isnan(cref.remap_columns(h0, f, h1b, PPM, wet)[:, 0, 0]).all()


@pytest.mark.parametrize("scheme", [PCM, PLM, PPM])
def test_target_means_are_integrals_of_the_reconstruction(scheme):
    # u_tgt h_tgt = the integral of the source cells' polynomials over the target cell (the thickest piece of each
    # source cell carries the cell's exact content, a round-off difference)
    rng = np.random.default_rng(21)
    n0, n1 = 16, 11
    h0 = rng.uniform(0.3, 3.0, n0)
    h0[[4, 9]] = 0.0
    u0 = np.cumsum(rng.normal(0, 1, n0)) + 10.0
    h1 = rng.uniform(0.2, 2.0, n1)
    h1 *= h0.sum() / h1.sum()
    active = (h0 > 0).sum()
    eff = PCM if active <= 1 else (PLM if active <= 3 and scheme > PLM else scheme)
    coef = cref.reconstruct(h0, u0, eff)
    z0 = np.concatenate([[0.0], np.cumsum(h0)])
    z1 = np.concatenate([[0.0], np.cumsum(h1)])

    def p(z):
        i = int(np.clip(np.searchsorted(z0, z, side="right") - 1, 0, n0 - 1))
        while h0[i] == 0.0:
            i -= 1
        xi = (z - z0[i]) / h0[i]
        return coef[i, 0] + xi * (coef[i, 1] + xi * coef[i, 2])
    u1 = _remap(h0, u0, h1, scheme)
    for j in range(n1):
        pts = [z for z in z0 if z1[j] < z < z1[j + 1]]
        val = integrate.quad(p, z1[j], z1[j + 1], points=pts or None, limit=200)[0]
        np.testing.assert_allclose(u1[j] * h1[j], val, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose((h1 * u1).sum(), (h0 * u0).sum(), rtol=1e-13)


# ------------------------------------------------------------------------------------------------ regrid
def test_zstar_full_depth_shallow_and_tiny_columns():
    H, eta = 201.0, 0.5
    zw = cref.zstar_column(H, eta, DELR, H_MIN)
    s = (H + eta) / H
    np.testing.assert_allclose(zw, (eta - s * np.r_[0.0, np.cumsum(DELR)])[::-1], rtol=1e-14, atol=1e-12)
    zw = cref.zstar_column(40.0, 0.0, DELR, H_MIN)
    h = np.diff(zw)
    assert zw[0] == -40.0 and zw[-1] == 0.0
    np.testing.assert_allclose(h.sum(), 40.0, rtol=1e-14)
    assert (h >= H_MIN - 1e-12).all()
    np.testing.assert_allclose(h[-4:][::-1], DELR[:4], rtol=1e-12)
    assert np.count_nonzero(np.isclose(h, H_MIN, rtol=0, atol=1e-12)) == 3
    np.testing.assert_allclose(h[3], 9.0 - 3 * H_MIN, rtol=1e-12)
    h = np.diff(cref.zstar_column(0.005, 0.0, DELR, H_MIN))
    np.testing.assert_allclose(h.sum(), 0.005, rtol=1e-12)
    assert (h > 0).all()


def test_sigma_reproduces_fractions_and_keeps_stretched_shallow_layers_positive():
    ds = np.array([0.1, 0.2, 0.3, 0.4])
    zw = cref.sigma_column(100.0, 2.0, ds, H_MIN)
    np.testing.assert_allclose(np.diff(zw), 102.0 * ds, rtol=1e-14)
    assert zw[0] == -100.0 and zw[-1] == 2.0
    ds = np.r_[np.full(5, 0.01), np.full(5, 0.19)]
    for D in (1e-4, 2e-3, 0.02, 0.5):
        h = np.diff(cref.sigma_column(D, 0.0, ds, H_MIN))
        assert (h > 0).all()
        np.testing.assert_allclose(h.sum(), D, rtol=1e-12)


def test_regrid_keeps_dry_columns_and_preserves_the_column_thickness():
    rng = np.random.default_rng(8)
    n, ny, nx = 10, 5, 6
    H = rng.uniform(5.0, 250.0, (ny, nx))
    eta = rng.uniform(-0.5, 0.5, (ny, nx))
    active = rng.random((ny, nx)) > 0.2
    h = rng.uniform(0.1, 10.0, (n, ny, nx))
    h[:, 0, 0] = 0.0
    hz, wet = cref.regrid(H, eta, active, h, "zstar", DELR)
    assert not wet[0, 0] and np.array_equal(wet, active & (H + eta > 0) & (h.sum(0) > 0))
    np.testing.assert_array_equal(hz[:, ~wet], h[:, ~wet])
    np.testing.assert_allclose(hz.sum(0)[wet], (H + eta)[wet], rtol=1e-13)
    hs, _ = cref.regrid(H, eta, active, h, "sigma", np.full(n, 1.0 / n))
    np.testing.assert_allclose(hs[:, wet], np.broadcast_to((H + eta)[wet] / n, (n, wet.sum())), rtol=1e-12)


# ------------------------------------------------------------------------------------------------ hybrid rule
def _linear_column(H=201.0, n=10, rho_top=1024.0, drho_dz=0.02):
    zw = np.linspace(-H, 0.0, n + 1)
    zc = 0.5 * (zw[1:] + zw[:-1])
    return np.diff(zw), rho_top - drho_dz * zc


def _hyb(H, eta, h, rho, targets, dz=DELR, max_h=None, keep=False):
    return cref.hycom1_column(H, eta, h, rho, targets, dz, max_h, H_MIN, keep)


def test_hybrid_interfaces_sit_on_the_target_densities_and_respect_the_zstar_floor():
    h, rho = _linear_column()
    z_t = -np.array([0.0, 30, 60, 80, 100, 120, 140, 160, 170, 178, 201.0])
    zw = _hyb(201.0, 0.0, h, rho, 1024.0 - 0.02 * z_t)
    np.testing.assert_allclose(zw[::-1][1:-1], z_t[1:-1], atol=1e-9)
    zw = _hyb(201.0, 0.0, h, rho, np.full(11, 1024.0))
    np.testing.assert_allclose(zw, cref.zstar_column(201.0, 0.0, DELR, H_MIN), atol=1e-9)


def test_hybrid_unstratified_out_of_range_and_nonmonotonic_columns():
    zw = _hyb(201.0, 0.0, np.full(10, 20.1), np.full(10, 1025.0), np.linspace(1024.0, 1027.0, 11))
    assert np.isfinite(zw).all() and (np.diff(zw) > 0).all()
    h, rho = _linear_column()
    zw = _hyb(201.0, 0.0, h, rho, np.linspace(1000.0, 1100.0, 11))
    assert (np.diff(zw) >= H_MIN * (1 - 1e-12)).all()
    np.testing.assert_allclose(np.diff(zw).sum(), 201.0, rtol=1e-14)
    rho2 = rho.copy()
    rho2[3] += 0.5
    targets = np.linspace(1024.0, 1028.0, 11)
    zw = _hyb(201.0, 0.0, h, rho2, targets)
    assert (np.diff(zw) > 0).all()
    rho_m = rho2.copy()
    for k in range(1, rho_m.size):
        rho_m[k] = min(rho_m[k], rho_m[k - 1])
    np.testing.assert_allclose(zw, _hyb(201.0, 0.0, h, rho_m, targets), rtol=0, atol=1e-12)


def test_hybrid_massless_bed_layer_and_maximum_thickness():
    n = 10
    targets = np.linspace(1019.0, 1026.0, n + 1)
    h = np.full(n, H_MIN)
    h[1], h[8], h[9] = 0.7, 19.3 - 7 * H_MIN - 0.01, 0.01
    rho = np.full(n, 1020.0)
    rho[1:8] = 1025.0
    zw = _hyb(20.0, 0.0, h, rho, targets, dz=np.full(n, 0.01))
    assert (-20.0 + H_MIN + 0.7) - zw[n - 2] < 0.1 * 0.7
    h, rho = _linear_column()
    zw = _hyb(201.0, 0.0, h, rho, np.r_[1024.0, np.full(9, 1028.0), 1028.1], max_h=np.full(10, 25.0))
    assert (np.diff(zw) <= 25.0 + 1e-9).all()
    np.testing.assert_allclose(np.diff(zw).sum(), 201.0, rtol=1e-14)


def _window_case(H, d_sigma=50.0, d_hybrid=100.0, n=10):
    zw_t = np.linspace(-H, 0.0, n + 1)
    zc = 0.5 * (zw_t[1:] + zw_t[:-1])
    h, rho = np.diff(zw_t), 1024.0 - 0.02 * zc
    dz = np.full(n, 2.0)
    targets = 1024.0 + 0.02 * H * np.linspace(0.0, 1.0, n + 1) ** 2
    hyb = cref.hycom1_column(H, 0.0, h, rho, targets, dz, None, H_MIN, False)
    sig = cref.sigma_column(H, 0.0, np.full(n, 1.0 / n), H_MIN)
    zs = cref.zstar_column(H, 0.0, dz, H_MIN)
    return hyb, sig, zs, cref.window_shallow_sigma(hyb, sig, zs, H, H, d_sigma, d_hybrid, H_MIN)


def test_window_is_sigma_shallow_hybrid_deep_and_keeps_fitting_isopycnals():
    hyb, sig, _, out = _window_case(40.0)
    np.testing.assert_array_equal(out, sig)
    hyb, sig, _, out = _window_case(120.0)
    np.testing.assert_array_equal(out, hyb)
    H, q = 75.0, 0.5
    hyb, sig, zs, out = _window_case(H)
    lo, hi = q * (-H) + (1 - q) * sig, q * zs + (1 - q) * sig
    assert ((out >= lo - 1e-12) & (out <= hi + 1e-12)).all()
    inside = (hyb > lo) & (hyb < hi)
    inside[[0, -1]] = False
    assert inside.sum() == 3 and (~inside[1:-1]).sum() == 6
    np.testing.assert_allclose(out[inside], hyb[inside], atol=1e-12)
    for Hc, other in ((50.0 + 1e-6, 0), (100.0 - 1e-6, 1)):
        hyb, sig, _, out = _window_case(Hc)
        assert np.abs(out - (sig if other == 0 else hyb)).max() < 1e-4


def test_regrid_hybrid_columns_match_single_columns_and_the_sigma_zone():
    hyb, sig, zs, out = _window_case(75.0)
    n = 10
    zw_t = np.linspace(-75.0, 0.0, n + 1)
    zc = 0.5 * (zw_t[1:] + zw_t[:-1])
    h3 = np.diff(zw_t)[:, None, None].repeat(2, 1)
    r3 = (1024.0 - 0.02 * zc)[:, None, None].repeat(2, 1)
    params = dict(dz_top=np.full(n, 2.0), rho_target=1024.0 + 0.02 * 75.0 * np.linspace(0.0, 1.0, n + 1) ** 2,
                  dsigma=np.full(n, 1.0 / n), d_sigma=50.0, d_hybrid=100.0, keep_isopycnal=False)
    h_new, wet = cref.regrid_hybrid(np.full((2, 1), 75.0), np.zeros((2, 1)), np.ones((2, 1), bool), h3, r3, params)
    np.testing.assert_allclose(h_new[:, 0, 0], np.diff(out), atol=1e-12)
    # flooded land (H < 0) in the sigma zone: pure sigma at any depth
    H = np.array([[-0.5, 10.0]])
    eta = np.array([[0.7, 0.0]])
    base = dict(dz_top=np.ones(10), rho_target=np.linspace(1024, 1026, 11), dsigma=np.full(10, 0.1), d_sigma=5.0,
                d_hybrid=20.0)
    h_new, _ = cref.regrid_hybrid(H, eta, np.ones((1, 2), bool), np.ones((10, 1, 2)), np.full((10, 1, 2), 1025.0),
                                  base)
    np.testing.assert_allclose(h_new[:, 0, 0], 0.02, rtol=1e-12)


def _consistent_column(n=10, H=200.0, seed=0):
    rng = np.random.default_rng(seed)
    targets = np.linspace(1020.0, 1030.0, n + 1)
    mid_top = 0.5 * (targets[:-1] + targets[1:])
    h = rng.uniform(0.5, 1.5, n)
    h *= H / h.sum()
    return h, mid_top[::-1].copy(), targets


def test_keep_isopycnal_leaves_consistent_layers_and_regrids_only_inconsistent_interfaces():
    h, rho, targets = _consistent_column()
    dz = np.full(10, 1e-6)
    zin = -200.0 + np.concatenate([[0.0], np.cumsum(h)])
    zw = _hyb(200.0, 0.0, h, rho, targets, dz=dz, keep=True)
    np.testing.assert_allclose(zw, zin, rtol=0, atol=1e-11)
    assert np.abs(_hyb(200.0, 0.0, h, rho, targets, dz=dz) - zw).max() > 1.0
    rho[4] = 1021.2
    p1m = _hyb(200.0, 0.0, h, rho, targets, dz=dz)
    zw = _hyb(200.0, 0.0, h, rho, targets, dz=dz, keep=True)
    other = [k for k in range(11) if k != 5]
    np.testing.assert_allclose(zw[other], zin[other], rtol=0, atol=1e-11)
    np.testing.assert_allclose(zw[5], np.clip(p1m[5], zin[4], zin[6]), rtol=0, atol=1e-11)


def test_keep_isopycnal_massless_layers_floor_and_jump():
    h, rho, targets = _consistent_column()
    h[3] = H_MIN
    rho[3] = 1029.9
    h *= 200.0 / h.sum()
    zw = _hyb(200.0, 0.0, h, rho, targets, dz=np.full(10, 1e-6), keep=True)
    np.testing.assert_allclose(zw, -200.0 + np.concatenate([[0.0], np.cumsum(h)]), rtol=0, atol=1e-11)
    h, rho, targets = _consistent_column()
    h[-3:] = [0.2, 0.2, 0.2]
    h *= 200.0 / h.sum()
    dz = np.full(10, 5.0)
    zw = _hyb(200.0, 0.0, h, rho, targets, dz=dz, keep=True)
    assert (np.diff(zw) > 0).all() and zw[0] == -200.0 and zw[-1] == 0.0
    np.testing.assert_allclose(zw[::-1][1:4], -np.cumsum(dz)[:3], atol=1e-11)
    h_top = np.array([10.0, 90.0, 10.0, 30.0, 30.0, 30.0])
    rho_top = np.array([1001.2, 1001.2, 1002.0, 1002.0, 1002.0, 1002.0])
    targets = np.array([999.5, 1001.0, 1001.5, 1001.8, 1001.9, 1001.95, 1001.99])
    dz = np.array([10.0, 40.0, 60.0, 30.0, 30.0, 30.0])
    zw = cref.hycom1_column(200.0, 0.0, h_top[::-1].copy(), rho_top[::-1].copy(), targets, dz, None, H_MIN, True)
    np.testing.assert_allclose(-zw[::-1], [0.0, 10.0, 100.0, 110.0, 140.0, 170.0, 200.0], atol=1e-9)
