"""
Pins of the oracle's UP3 / UP1 faces and RHS (oracle.numpy_ref.up_faces, faces_up_h, rhs_up; the reference's
_up3_faces, sigma_baroclinic.py:164-187, and rhs_numpy, tracers.py:139-151).  The reference's tests hold no vectors
for the tracer upwind faces, so these checks come from what the mathematics fixes: UP3 reproduces the face value of
any quadratic from its cell averages, each branch reads only its upwind-biased stencil, a cell that is not ok drops
the face to the donor value, UP1 is the donor flux of upwind_faces_h, the flux form telescopes, and a linear RHS
advanced by SSP-RK3 is the third-order Taylor polynomial of the step.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import numpy_ref as R


def _quadratic_averages(n, a, b, c):
    """Cell averages over [m, m+1] of p(x) = a + b x + c x^2, and p itself."""
    m = np.arange(n, dtype=float)
    avg = a + b * (m + 0.5) + c * ((m + 1.0) ** 3 - m ** 3) / 3.0
    return avg, lambda x: a + b * x + c * x * x


@pytest.mark.parametrize("axis", [1, 2, 3])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_up3_reproduces_the_face_values_of_a_quadratic(axis, sign):
    n = 9
    avg, p = _quadratic_averages(n, 0.7, -1.3, 0.45)
    shape = [1, 1, 1, 1]
    shape[axis] = n
    c = avg.reshape(shape)
    ok = np.ones(c.shape, bool)
    fshape = list(c.shape)
    fshape[axis] = n + 1
    flux = np.full(fshape, sign)
    face = np.moveaxis(R.up_faces(c, flux, ok, axis, order=3), axis, -1).ravel()
    f = np.arange(1, n)
    # faces 1 and n-1 lack the outer neighbour of one branch and fall back to the donor cell
    inner = f[(f >= 2) if sign > 0 else (f <= n - 2)]
    np.testing.assert_allclose(face[inner], p(inner.astype(float)), rtol=0, atol=1e-13)
    assert face[0] == 0.0 and face[n] == 0.0
    edge = 1 if sign > 0 else n - 1
    assert face[edge] == (avg[edge - 1] if sign > 0 else avg[edge])


def test_each_branch_reads_only_its_upwind_stencil():
    rng = np.random.default_rng(3)
    n = 10
    c = rng.normal(size=(1, 1, 1, n))
    ok = np.ones(c.shape, bool)
    for sign, far, unused in ((1.0, -2, 1), (-1.0, 1, -2)):
        q = np.full((1, 1, 1, n + 1), sign)
        base = R.up_faces(c, q, ok, 3)[..., 5]
        moved = c.copy()
        moved[..., 5 + unused] += 10.0          # the cell the branch must not read
        assert R.up_faces(moved, q, ok, 3)[..., 5] == base
        moved = c.copy()
        moved[..., 5 + far] += 10.0             # its outer stencil cell
        assert R.up_faces(moved, q, ok, 3)[..., 5] != base


def test_a_cell_that_is_not_ok_drops_the_face_to_the_donor_value():
    rng = np.random.default_rng(4)
    n = 8
    c = rng.normal(size=(1, 1, 1, n))
    ok = np.ones(c.shape, bool)
    ok[..., 2] = False
    qp = np.full((1, 1, 1, n + 1), 1.0)
    qn = -qp
    fp = R.up_faces(c, qp, ok, 3).ravel()
    fn = R.up_faces(c, qn, ok, 3).ravel()
    donor_p = R.up_faces(c, qp, ok, 3, order=1).ravel()
    donor_n = R.up_faces(c, qn, ok, 3, order=1).ravel()
    for f in (2, 3, 4):                          # stencils (f-2 .. f) that hold cell 2
        assert fp[f] == donor_p[f] == c.ravel()[f - 1]
    for f in (1, 2, 3):                          # stencils (f-1 .. f+1) that hold cell 2
        assert fn[f] == donor_n[f] == c.ravel()[f]
    assert fp[5] != donor_p[5] and fn[5] != donor_n[5]


def test_up1_interior_fluxes_equal_the_donor_cell_fluxes():
    rng = np.random.default_rng(5)
    nt, n, ny, nx = 2, 3, 7, 9
    c = rng.normal(size=(nt, n, ny, nx))
    Qx = rng.normal(size=(n, ny, nx + 1))
    Qy = rng.normal(size=(n, ny + 1, nx))
    ok = rng.random((n, ny, nx)) > 0.2
    for mask in (0, 15):
        fx, fy = R.faces_up_h(c, Qx, Qy, ok, c + 1.0, mask, order=1)
        gx, gy = R.upwind_faces_h(c, Qx, Qy, c + 1.0, mask)
        assert np.array_equal(fx[..., 1:-1], gx[..., 1:-1]) and np.array_equal(fy[..., 1:-1, :], gy[..., 1:-1, :])
        if mask:
            assert np.array_equal(fx, gx) and np.array_equal(fy, gy)


@pytest.mark.parametrize("order", [3, 1])
def test_closed_box_rhs_telescopes_to_zero(order):
    rng = np.random.default_rng(6)
    nt, n, ny, nx = 2, 5, 8, 11
    c = rng.normal(size=(nt, n, ny, nx))
    Qx = rng.normal(size=(n, ny, nx + 1))
    Qy = rng.normal(size=(n, ny + 1, nx))
    Fz = rng.normal(size=(n + 1, ny, nx))
    ok = rng.random((n, ny, nx)) > 0.1
    r = R.rhs_up(c, Qx, Qy, Fz, ok, c, 0, order)
    scale = np.abs(c).max() * (np.abs(Qx).sum() + np.abs(Qy).sum() + np.abs(Fz).sum()) / nt
    assert np.all(np.abs(r.sum(axis=(1, 2, 3))) < 1e-13 * scale)
    # the bed face is Fz[0] * 0 and the surface face 0: Fz[0] and Fz[n] do not enter
    Fz2 = Fz.copy()
    Fz2[0] = 7.0
    Fz2[n] = -3.0
    assert np.array_equal(R.rhs_up(c, Qx, Qy, Fz2, ok, c, 0, order), r)


@pytest.mark.parametrize("order", [3, 1])
def test_uniform_tracer_in_a_uniform_through_flow_is_steady(order):
    n, ny, nx = 2, 4, 9
    c = np.full((1, n, ny, nx), 3.25)
    Qx = np.full((n, ny, nx + 1), 0.4)
    Qy = np.zeros((n, ny + 1, nx))
    Fz = np.zeros((n + 1, ny, nx))
    r = R.rhs_up(c, Qx, Qy, Fz, np.ones((n, ny, nx), bool), c, R.OPEN_W | R.OPEN_E, order)
    assert np.all(r == 0.0)


@pytest.mark.parametrize("order", [3, 1])
def test_rk3_step_is_the_cubic_taylor_polynomial_of_the_linear_rhs(order):
    """With V0 = V1 = 1 and a linear RHS r = A c, SSP-RK3 gives (I + dt A + (dt A)^2/2 + (dt A)^3/6) c."""
    rng = np.random.default_rng(8)
    n, ny, nx = 3, 2, 7
    Qx = rng.normal(size=(n, ny, nx + 1))
    Qy = rng.normal(size=(n, ny + 1, nx))
    Fz = rng.normal(size=(n + 1, ny, nx))
    ok = np.ones((n, ny, nx), bool)
    m = n * ny * nx
    A = np.empty((m, m))
    for j in range(m):
        e = np.zeros(m)
        e[j] = 1.0
        A[:, j] = R.rhs_up(e.reshape(1, n, ny, nx), Qx, Qy, Fz, ok, None, 0, order).ravel()
    c = rng.normal(size=(1, n, ny, nx))
    dt = 0.05
    V = np.ones((n, ny, nx))
    out = R.advect_rk3(c, V, V, Qx, Qy, Fz, ok, dt, up_order=order)
    B = dt * A
    want = (np.eye(m) + B + B @ B / 2.0 + B @ B @ B / 6.0) @ c.ravel()
    np.testing.assert_allclose(out.ravel(), want, rtol=0, atol=1e-12)
