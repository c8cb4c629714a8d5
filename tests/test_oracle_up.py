"""
Pins of the oracle's UP3 / UP1 and centred faces, their RHS and the 'pc' step (oracle.numpy_ref.up_faces,
centred_faces, faces_up_h, rhs_up, advect_pc; the reference's _up3_faces and _centred_faces,
sigma_baroclinic.py:164-187, 632-652, rhs_numpy and advect_numpy, tracers.py:139-151, 303-330).  The reference's
tests hold no vectors for these faces; the checks are its sixth-order test of the centred faces
(test_sigma_baroclinic.py:133-143) and what the mathematics fixes: UP3 reproduces the face value of any quadratic
from its cell averages and the centred faces that of a quintic (cubic, linear next to walls), each UP3 branch reads
only its upwind-biased stencil, a cell that is not ok drops the face to the donor value, UP1 is the donor flux of
upwind_faces_h, the flux form telescopes, and a linear RHS advanced by SSP-RK3 is the third-order Taylor polynomial
of the step (by 'pc', the predictor-corrector product).
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


# ------------------------------------------------------------------------------------------------ centred faces
def _poly_averages(n, coef):
    """Cell averages over [m, m+1] of the polynomial sum coef[p] x^p, and the polynomial."""
    m = np.arange(n, dtype=float)
    avg = sum(c * ((m + 1.0) ** (p + 1) - m ** (p + 1)) / (p + 1) for p, c in enumerate(coef))
    return avg, lambda x: sum(c * x ** p for p, c in enumerate(coef))


def test_centred_faces_are_sixth_order():
    """The reference's own test (test_sigma_baroclinic.py:133-143) on the oracle."""
    errs = []
    for n in (40, 80):
        x = (np.arange(n) + 0.5) / n
        h = 1.0 / n
        c = (np.cos(2 * np.pi * (x - h / 2)) - np.cos(2 * np.pi * (x + h / 2))) / (2 * np.pi * h)
        faces = R.centred_faces(c, np.ones(n, bool), axis=0)
        exact = np.sin(2 * np.pi * np.arange(n + 1) / n)
        errs.append(np.abs(faces[4:-3] - exact[4:-3]).max())
    assert np.log2(errs[0] / errs[1]) > 5.0


@pytest.mark.parametrize("axis", [1, 2, 3])
def test_centred_faces_reproduce_polynomials_by_stencil_width(axis):
    n = 12
    shape = [1, 1, 1, 1]
    shape[axis] = n
    ok = np.ones(shape, bool)
    f = np.arange(n + 1, dtype=float)
    for coef, faces in (([0.3, -1.1, 0.7, 0.25, -0.05, 0.004], range(3, n - 2)),   # 6-cell stencil: quintic
                        ([0.3, -1.1, 0.7, 0.25], (2, n - 2)),                      # 4-cell stencil: cubic
                        ([0.3, -1.1], (1, n - 1))):                                # 2 cells: linear
        avg, p = _poly_averages(n, coef)
        got = np.moveaxis(R.centred_faces(avg.reshape(shape), ok, axis), axis, -1).ravel()
        idx = np.array(list(faces))
        np.testing.assert_allclose(got[idx], p(f[idx]), rtol=0, atol=1e-11)
        assert got[0] == 0.0 and got[n] == 0.0


def test_centred_faces_ignore_the_flux_and_narrow_at_cells_that_are_not_ok():
    rng = np.random.default_rng(10)
    n = 12
    c = rng.normal(size=(1, 1, 1, n))
    ok = np.ones(c.shape, bool)
    q = rng.normal(size=(1, 1, 1, n + 1))
    a = R.faces_up_h(c, q[0], np.zeros((1, 2, n)), ok, c, 0, "c")[0] / q
    assert np.allclose(a[..., 1:-1], R.centred_faces(c, ok, 3)[..., 1:-1], rtol=1e-15, atol=0)
    ok[..., 6] = False
    wide = R.centred_faces(c, np.ones(c.shape, bool), 3).ravel()
    got = R.centred_faces(c, ok, 3).ravel()
    cv = c.ravel()
    assert got[4] == (7.0 * (cv[3] + cv[4]) - (cv[2] + cv[5])) / 12.0      # 6-stencil holds cell 6: 4-point
    assert got[6] == 0.5 * (cv[5] + cv[6]) and got[7] == 0.5 * (cv[6] + cv[7])
    assert got[10] == wide[10] and got[3] == wide[3]


def test_centred_rhs_telescopes_to_zero():
    rng = np.random.default_rng(11)
    nt, n, ny, nx = 2, 6, 8, 11
    c = rng.normal(size=(nt, n, ny, nx))
    Qx, Qy, Fz = rng.normal(size=(n, ny, nx + 1)), rng.normal(size=(n, ny + 1, nx)), rng.normal(size=(n + 1, ny, nx))
    r = R.rhs_up(c, Qx, Qy, Fz, rng.random((n, ny, nx)) > 0.1, c, 0, "c")
    scale = np.abs(c).max() * (np.abs(Qx).sum() + np.abs(Qy).sum() + np.abs(Fz).sum()) / nt
    assert np.all(np.abs(r.sum(axis=(1, 2, 3))) < 1e-13 * scale)


@pytest.mark.parametrize("up_order", [0, 3, 1])
def test_pc_step_of_a_linear_rhs_is_the_predictor_corrector_product(up_order):
    """With V0 = V1 = 1 and linear RHS A_c (centred) and A (the corrector's faces): out = (I + dt A (I + dt/2 A_c)) c.
    WENO is linear in c only for a fixed stencil choice, so its case uses a constant tracer plus a small ripple
    and compares against the matrix of the WENO RHS linearised by finite differences (loose)."""
    rng = np.random.default_rng(12)
    n, ny, nx = 3, 2, 9
    Qx, Qy, Fz = rng.normal(size=(n, ny, nx + 1)), rng.normal(size=(n, ny + 1, nx)), rng.normal(size=(n + 1, ny, nx))
    ok = np.ones((n, ny, nx), bool)
    V = np.ones((n, ny, nx))
    dt = 0.05
    m = n * ny * nx

    def matrix(op):
        A = np.empty((m, m))
        for j in range(m):
            e = np.zeros(m)
            e[j] = 1.0
            A[:, j] = op(e.reshape(1, n, ny, nx)).ravel()
        return A
    Ac = matrix(lambda e: R.rhs_up(e, Qx, Qy, Fz, ok, None, 0, "c"))
    c = rng.normal(size=(1, n, ny, nx))
    if up_order:
        A = matrix(lambda e: R.rhs_up(e, Qx, Qy, Fz, ok, None, 0, up_order))
        want = (np.eye(m) + dt * A @ (np.eye(m) + 0.5 * dt * Ac)) @ c.ravel()
        out = R.advect_pc(c, V, V, Qx, Qy, Fz, ok, dt, up_order=up_order)
        np.testing.assert_allclose(out.ravel(), want, rtol=0, atol=1e-12)
    else:
        # the WENO corrector applied to the exact predictor state
        ch = (c.ravel() + 0.5 * dt * Ac @ c.ravel()).reshape(c.shape)
        want = c + dt * R.rhs(ch, Qx, Qy, Fz, ok, ch, 0, True, True)
        out = R.advect_pc(c, V, V, Qx, Qy, Fz, ok, dt)
        np.testing.assert_allclose(out, want, rtol=0, atol=1e-12)
