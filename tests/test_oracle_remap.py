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
