"""
Pins of the oracle's vertical WENO pass (CPU): ``numpy_ref.faces_z``, ``rhs`` with ``vert`` and the SSP-RK3 step
with Fz != 0, in both restatements (NumPy and the C loops).  The reference's vertical pass
(core/dynamics/sigma_baroclinic.py:335-353) forms, on interface k between layers k-1 and k (k = 1 .. N-1), the
WENO value of the upwind-ordered layers (k-3 .. k+1 for Fz[k] >= 0, k+2 .. k-2 for Fz[k] < 0) times Fz[k];
Fz == 0 is skipped, the bed and surface faces are 0.  What fixes each part:

* the interface index and the flux index: cell means of polynomials along k give the polynomial's value at
  z = k (not k +- 1), times Fz[k] (Fz varies with k, so Fz[k +- 1] fails);
* the upwind side: at a step the face takes the value of the side the flux comes from (an upwind-side flip gives
  the other side's value), for both signs;
* the stencil: fifth order along k (log2 error ratio > 4, test_sigma_baroclinic.py:101-111) and no overshoot at a
  step (:114-119), the WENO3 / donor fallbacks next to the bed, the surface and a dry layer;
* the whole step: a uniform tracer stays uniform to 1e-11 with content kept to 1e-13 over 200 steps under a
  baroclinic exchange flow whose continuity Fz is cumulated from the bed (test_sigma_baroclinic.py:28-40 with
  layer_transport.py:85-88), and a bump carried up a tall column by a uniform Fz moves up by the Courant number
  per step.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import numpy_ref as R


def _faces_numpy(c, Fz, ok, z=True):
    return R.faces_z(c, Fz, ok, z)


def _faces_c(c, Fz, ok, z=True):
    """z faces of the C loops, from their RHS with no horizontal transport: rhs_k = -(fz[k+1] - fz[k]), fz[0] = 0."""
    from oracle import cref
    nt, n, ny, nx = c.shape
    Qx = np.zeros((n, ny, nx + 1))
    Qy = np.zeros((n, ny + 1, nx))
    r = cref.rhs(c, Qx, Qy, Fz, ok, c, 0, z, True)
    fz = np.zeros((nt, n + 1, ny, nx))
    fz[:, 1:] = -np.cumsum(r, axis=1)
    return fz


FACES = {"numpy": _faces_numpy, "c": _faces_c}


def _column(values, ny=2, nx=3):
    """A stack (1, n, ny, nx) whose every column holds ``values`` along k."""
    v = np.asarray(values, float)
    return np.broadcast_to(v[None, :, None, None], (1, v.size, ny, nx)).copy()


def _flux(n, sign, ny=2, nx=3):
    """Fz (n+1, ny, nx) of one sign, different on every interface (so a shifted flux index shows), nonzero at the
    bed and the surface too (which must be ignored)."""
    f = sign * (1.0 + 0.25 * np.arange(n + 1))
    return np.broadcast_to(f[:, None, None], (n + 1, ny, nx)).copy()


def _quad_means(n, a=1.5, b=-0.4, q=0.03):
    """Cell means over [k, k+1] of p(z) = a + b z + q z^2, and p itself."""
    k = np.arange(n, dtype=float)
    means = a + b * (k + 0.5) + q * (k * k + k + 1.0 / 3.0)
    return means, (lambda z: a + b * z + q * z * z)


@pytest.mark.parametrize("impl", list(FACES))
@pytest.mark.parametrize("sign", [1.0, -1.0])
@pytest.mark.parametrize("z", [True, False])
def test_quadratic_cell_means_give_the_interface_value_times_its_own_flux(impl, sign, z):
    n = 14
    means, p = _quad_means(n)
    c = _column(means)
    Fz = _flux(n, sign)
    ok = np.ones((n, 2, 3), bool)
    fz = FACES[impl](c, Fz, ok, z)
    assert np.all(fz[:, 0] == 0.0) and np.all(fz[:, n] == 0.0)          # bed and surface closed
    # interfaces whose five-cell upwind stencil lies inside the column: exact value p(k), times Fz[k]
    ks = range(3, n - 1) if sign > 0 else range(2, n - 2)
    for k in ks:
        np.testing.assert_allclose(fz[0, k], Fz[k] * p(float(k)), rtol=1e-12, atol=0)
    # interface 1: for Fz >= 0 the stencil v2 = layer -1 is missing, so the donor layer 0 is taken (no WENO3);
    # for Fz < 0 the stencil k+2 .. k-2 = 3 .. -1 has v2 .. v4 = 2, 1, 0, so WENO3, exact only for linear data
    if sign > 0:
        np.testing.assert_allclose(fz[0, 1], Fz[1] * means[0], rtol=1e-14)
    # the surface side: interface n-1 for Fz < 0 takes the donor layer n-1 (v2 = layer n+1 is missing)
    if sign < 0:
        np.testing.assert_allclose(fz[0, n - 1], Fz[n - 1] * means[n - 1], rtol=1e-14)


@pytest.mark.parametrize("impl", list(FACES))
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_linear_cell_means_are_exact_on_the_weno3_faces_too(impl, sign):
    n = 10
    k = np.arange(n, dtype=float)
    c = _column(2.0 + 0.3 * (k + 0.5))
    Fz = _flux(n, sign)
    fz = FACES[impl](c, Fz, np.ones((n, 2, 3), bool))
    # WENO5 where the five cells exist, WENO3 on interface 2 (Fz > 0: v2..v4 = layers 0..2) and n-2 (Fz < 0)
    ks = range(2, n - 1) if sign > 0 else range(1, n - 2)
    for kk in ks:
        np.testing.assert_allclose(fz[0, kk], Fz[kk] * (2.0 + 0.3 * kk), rtol=1e-13)


@pytest.mark.parametrize("impl", list(FACES))
def test_a_step_takes_the_value_of_the_side_the_flux_comes_from(impl):
    # layers 0..7 hold 0, layers 8..15 hold 1; interface 8 lies on the step.  Upward flux (Fz > 0) brings layer 7's
    # water, downward flux layer 8's; WENO keeps each side's value to within its non-oscillatory tolerance
    n, m = 16, 8
    c = _column(np.where(np.arange(n) < m, 0.0, 1.0))
    ok = np.ones((n, 2, 3), bool)
    for sign, want in ((1.0, 0.0), (-1.0, 1.0)):
        Fz = _flux(n, sign)
        fz = FACES[impl](c, Fz, ok)
        v = fz[0, m] / Fz[m]
        assert np.abs(v - want).max() < 1e-6, (sign, v)
        # one interface below / above the step the upwind stencil straddles it: still no overshoot
        vals = fz[0, 1:n] / Fz[1:n]
        assert vals.min() > -1e-3 and vals.max() < 1.0 + 1e-3


@pytest.mark.parametrize("impl", list(FACES))
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_faces_are_fifth_order_along_the_layers(impl, sign):
    # test_sigma_baroclinic.py:101-111 on the layer axis: cell means of sin(2 pi z / n) over n = 40 and 80 layers
    errs = []
    for n in (40, 80):
        k = np.arange(n, dtype=float)
        h = 1.0 / n
        x = (k + 0.5) / n
        means = (np.cos(2 * np.pi * (x - h / 2)) - np.cos(2 * np.pi * (x + h / 2))) / (2 * np.pi * h)
        c = _column(means, 1, 1)
        Fz = np.full((n + 1, 1, 1), sign)
        fz = FACES[impl](c, Fz, np.ones((n, 1, 1), bool))
        exact = np.sin(2 * np.pi * np.arange(n + 1) / n)
        errs.append(np.abs(fz[0, 4:-3, 0, 0] / sign - exact[4:-3]).max())
    assert np.log2(errs[0] / errs[1]) > 4.0


@pytest.mark.parametrize("impl", list(FACES))
def test_a_dry_layer_cuts_the_stencil_of_the_faces_that_hold_it(impl):
    # linear data with layer 6 not ok: interfaces whose five-cell stencil holds layer 6 but whose v2..v4 do not
    # fall back to WENO3 (still exact for linear data); those whose v2..v4 hold it take the donor layer
    n, dry = 14, 6
    k = np.arange(n, dtype=float)
    means = 1.0 + 0.5 * (k + 0.5)
    c = _column(means)
    ok = np.ones((n, 2, 3), bool)
    ok[dry] = False
    Fz = _flux(n, 1.0)                                       # upward: stencil k-3 .. k+1, v2..v4 = k-2 .. k
    fz = FACES[impl](c, Fz, ok)
    for kk in range(3, n - 1):
        stencil = range(kk - 3, kk + 2)
        inner = range(kk - 2, kk + 1)
        v = fz[0, kk, 0, 0] / Fz[kk, 0, 0]
        if dry not in stencil:
            want = 1.0 + 0.5 * kk                            # WENO5, exact
        elif dry not in inner:
            want = 1.0 + 0.5 * kk                            # WENO3 on k-2 .. k, exact for linear data
        else:
            want = means[kk - 1]                             # the donor layer below
        assert abs(v - want) < 1e-12 * max(1.0, abs(want)), (kk, v, want)


@pytest.mark.parametrize("impl", list(FACES))
def test_zero_flux_interfaces_are_skipped(impl):
    # Fz == 0 contributes nothing, even next to a non-finite value (the reference skips the face, :341-342)
    n = 8
    c = _column(np.linspace(1.0, 2.0, n))
    c[0, 4] = np.inf
    Fz = _flux(n, 1.0)
    Fz[3:7] = 0.0                                            # every interface whose stencil holds layer 4 ...
    Fz[1:3] = 0.0                                            # ... and the ones below it
    Fz[7] = 0.0
    fz = FACES[impl](c, Fz, np.ones((n, 2, 3), bool))
    assert np.all(fz == 0.0)


def _exchange_case(nx=18, ny=14, nz=8, seed=3):
    """V1 = V0 under a baroclinic exchange flow: layer k carries a_k (Ux, Uy) with sum_k a_k = 0 and (Ux, Uy)
    horizontally divergent, so each layer's divergence leaves through the interfaces, Fz cumulated from the bed
    (layer_transport.py:85-88), and the surface flux vanishes.  Closed edges."""
    rng = np.random.default_rng(seed)
    a = np.cos(np.pi * (np.arange(nz) + 0.5) / nz)
    a -= a.mean()
    Ux = rng.normal(0.0, 40.0, (ny, nx + 1))
    Uy = rng.normal(0.0, 40.0, (ny + 1, nx))
    Ux[:, 0] = Ux[:, -1] = 0.0
    Uy[0] = Uy[-1] = 0.0
    Qx = a[:, None, None] * Ux[None]
    Qy = a[:, None, None] * Uy[None]
    div = Qx[..., 1:] - Qx[..., :-1] + Qy[:, 1:] - Qy[:, :-1]
    Fz = np.zeros((nz + 1, ny, nx))
    Fz[1:] = -np.cumsum(div, axis=0)
    Fz[nz] = 0.0
    V = rng.uniform(0.8e6, 1.2e6, (nz, ny, nx))
    ok = np.ones((nz, ny, nx), bool)
    return V, Qx, Qy, Fz, ok


@pytest.mark.parametrize("impl", ["numpy", "c"])
def test_uniform_tracer_stays_uniform_and_is_conserved_under_the_continuity_flux(impl, oracle_c):
    # test_sigma_baroclinic.py:28-40: a uniform tracer stays uniform to 1e-11 with content kept to 1e-13 over 200
    # steps while the interfaces carry the flow's continuity residual
    V, Qx, Qy, Fz, ok = _exchange_case()
    dt = 0.3 * V.min() / (np.abs(Qx).max() + np.abs(Qy).max() + np.abs(Fz).max()) / 2.0
    assert np.abs(Fz[1:-1]).max() > 0.1 * np.abs(Qx).max()         # the vertical pass really works
    step = oracle_c.advect_rk3 if impl == "c" else R.advect_rk3
    C = np.full((1,) + V.shape, 7.0)
    c0 = (V * C).sum()
    for _ in range(200 if impl == "c" else 40):
        C = step(C, V, V, Qx, Qy, Fz, ok, dt)
    np.testing.assert_allclose(C, 7.0, rtol=0, atol=1e-11)
    assert abs((V * C).sum() / c0 - 1.0) < 1e-13


def test_a_wrong_flux_index_breaks_uniformity():
    # the check above would fail if the vertical pass took Fz[k+1] on interface k: planted in a copy of the oracle
    V, Qx, Qy, Fz, ok = _exchange_case()
    c = np.full((1,) + V.shape, 7.0)
    fz = R.faces_z(c, Fz, ok)
    shifted = np.zeros_like(fz)
    shifted[:, 1:-1] = 7.0 * Fz[None, 2:]
    div = Qx[..., 1:] - Qx[..., :-1] + Qy[:, 1:] - Qy[:, :-1]
    good = -(div * 7.0 + fz[:, 1:] - fz[:, :-1])
    bad = -(div * 7.0 + shifted[:, 1:] - shifted[:, :-1])
    assert np.abs(good).max() < 1e-9 * np.abs(div * 7.0).max()
    assert np.abs(bad).max() > 1e-3 * np.abs(div * 7.0).max()


@pytest.mark.parametrize("impl", ["numpy", "c"])
def test_a_bump_moves_up_a_tall_column_at_the_courant_number(impl, oracle_c):
    # uniform interior Fz > 0 in a column of 160 layers of equal volume: the bed layer empties into the column and
    # the top layer fills (V1 from continuity), far from the bump.  40 steps at Courant 0.4 move it 16 layers up.
    n, nu, steps = 160, 0.4, 40
    V0 = np.full((n, 1, 1), 1.0)
    Fz = np.zeros((n + 1, 1, 1))
    Fz[1:n] = nu
    dt = 1.0
    V1 = V0 - dt * (Fz[1:] - Fz[:-1])
    k = np.arange(n) + 0.5
    bump = lambda s: np.exp(-0.5 * ((k - 50.0 - s) / 6.0) ** 2)          # noqa: E731
    C = bump(0.0)[None, :, None, None].copy()
    Qx = np.zeros((n, 1, 2))
    Qy = np.zeros((n, 2, 1))
    ok = np.ones((n, 1, 1), bool)
    step = oracle_c.advect_rk3 if impl == "c" else R.advect_rk3
    for _ in range(steps):
        C = step(C, V0, V1, Qx, Qy, Fz, ok, dt)
    c = C[0, :, 0, 0]
    assert int(np.argmax(c)) == int(np.argmax(bump(nu * steps)))
    assert np.abs(c - bump(nu * steps)).max() < 2e-3
    assert c.min() > -1e-6
