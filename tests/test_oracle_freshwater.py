"""
The oracle's freshwater substep, pinned by the reference's unit case (tests/test_n2a_freshwater.py:195-222: the
evaporation-deficit count, the column sums, the floor of the top layer, T unchanged, the content of S and a passive
tracer kept, the salinity-variance source) and by closed forms: rain touches the top layer only and scales the
salt-like members by V_adv / V_fw; a column holding less than its floors is emptied to zero.
"""
import numpy as np
import pytest

from oracle import numpy_ref as R


def _unit():
    rng = np.random.default_rng(1)
    n, ny, nx = 4, 4, 5
    lay = {"T": 0, "T2": 1, "S": 2, "tr": 3, "n": 4}
    T = 10 + rng.random((n, ny, nx))
    S = 30 + rng.random((n, ny, nx))
    stack = np.stack([T, T**2, S, rng.random((n, ny, nx))])
    V_adv = 1.0 + rng.random((n, ny, nx))
    dV = np.zeros((ny, nx))
    dV[0, :] = 0.2
    dV[1, :] = -0.1
    dV[2, 0] = -(V_adv[-1, 2, 0] + 0.5)
    return lay, T, S, stack, V_adv, dV


def test_reference_unit_case():
    lay, T, S, stack, V_adv, dV = _unit()
    V_fw, n_lim = R.freshwater_volumes(V_adv, dV, np.full(dV.shape, 1e-6))
    assert n_lim == 1
    np.testing.assert_allclose(V_fw.sum(0), V_adv.sum(0) + dV, rtol=1e-14)
    assert V_fw[-1, 2, 0] == pytest.approx(1e-6) and V_fw[-2, 2, 0] < V_adv[-2, 2, 0]
    out, n2, var_surf = R.apply_freshwater(stack.copy(), lay, V_adv, V_fw, dV, dt=10.0)
    assert n2 == 1
    np.testing.assert_array_equal(out[0], T)
    np.testing.assert_array_equal(out[1], T**2)
    np.testing.assert_allclose((V_fw * out[2]).sum(), (V_adv * S).sum(), rtol=1e-14)
    np.testing.assert_allclose((V_fw * out[3]).sum(), (V_adv * stack[3]).sum(), rtol=1e-14)
    np.testing.assert_allclose(var_surf, (V_fw * out[2] ** 2 - V_adv * S**2) / 10.0, rtol=1e-14, atol=1e-14)
    assert np.all(var_surf[:, 3:, :] == 0)


def test_rain_touches_the_top_layer_only_and_dilutes_salt_like_members():
    lay, T, S, stack, V_adv, _ = _unit()
    dV = np.full(V_adv.shape[1:], 0.3)
    V_fw, n_lim = R.freshwater_volumes(V_adv, dV, 1e-6)
    assert n_lim == 0
    np.testing.assert_array_equal(V_fw[:-1], V_adv[:-1])
    np.testing.assert_array_equal(V_fw[-1], V_adv[-1] + 0.3)
    out, n2, _ = R.apply_freshwater(stack.copy(), lay, V_adv, V_fw, dV, dt=10.0)
    assert n2 == 0
    np.testing.assert_array_equal(out[2, :-1], S[:-1])
    np.testing.assert_array_equal(out[2, -1], S[-1] * (V_adv[-1] / V_fw[-1]))


def test_a_column_short_of_its_floors_is_emptied_and_the_floor_may_be_an_array():
    V = np.full((3, 1, 2), 0.5)
    dV = np.array([[-1.4, -2.0]])
    out, n_lim = R.freshwater_volumes(V, dV, np.array([[0.1, 0.1]]))
    assert n_lim == 2
    np.testing.assert_allclose(out[:, 0, 0], [0.1, 0.0, 0.0], atol=1e-15)        # floors first, then the top down
    np.testing.assert_array_equal(out[:, 0, 1], [0.0, 0.0, 0.0])
    out_s, _ = R.freshwater_volumes(V, dV, 0.1)
    np.testing.assert_array_equal(out_s, out)
