"""
The oracle's layer densities of the hybrid regrid rule, pinned by the reference's own check values: Wright's (1997)
full-range density equals compiled MOM6 WRIGHT_FULL on the 2048 points of the reference's test data
(tests/test_eos_nonlinear.py:33-60, rtol 1e-12; the fixture tests/golden/mom6_eos_wright.npz), and its coefficients
equal the compiled ones bit for bit; the linear EOS by its closed forms (tests/test_eos.py:99-121, 222-227).  Also the
host side of the device EOS (names, defaults, errors), with no GPU call.
"""
from pathlib import Path

import numpy as np
import pytest

from oracle import numpy_ref as R

G = np.load(Path(__file__).parent / "golden" / "mom6_eos_wright.npz")


def test_wright_density_equals_compiled_mom6():
    rho = R.wright_density(G["T"], G["S"], G["p"])
    np.testing.assert_allclose(rho, G["w_rho"], rtol=1e-12, atol=0)
    assert G["T"].size == 2048 and np.ptp(G["p"]) > 5e7 and np.ptp(G["T"]) > 30.0


def test_wright_coefficients_equal_the_compiled_values():
    names = [str(n) for n in G["coef_names"]]
    assert len(names) == 15
    assert all(R.WRIGHT[n] == v for n, v in zip(names, G["coef_values"]))


def test_wright_limits_and_monotonicity():
    # at p = 0, 0 degC, S = 0: rho = b0 / (c0 + a0 b0) (the closed form of the fit)
    w = R.WRIGHT
    assert R.wright_density(0.0, 0.0, 0.0) == w["b0"] / (w["c0"] + w["a0"] * w["b0"])
    T = np.linspace(5.0, 30.0, 26)
    assert (np.diff(R.wright_density(T, 35.0, 0.0)) < 0).all()                    # warmer is lighter
    assert (np.diff(R.wright_density(10.0, np.linspace(30, 40, 11), 0.0)) > 0).all()   # saltier is denser
    assert R.wright_density(10.0, 35.0, 2e7) > R.wright_density(10.0, 35.0, 0.0) + 8.0     # compressible


def test_linear_density_closed_forms():
    assert R.linear_density(10.0, 35.0, 1025.0, 2e-4 * 1025, 10.0, 7.4e-4 * 1025, 35.0) == 1025.0
    assert R.linear_density(10.0) == 998.0
    lo, hi, mid = (R.linear_density(t, 35.0, s_coef=0.8) for t in (5.0, 15.0, 10.0))
    assert mid == pytest.approx(0.5 * (lo + hi), rel=1e-12)
    lo, hi, mid = (R.linear_density(10.0, s, s_coef=0.8) for s in (30.0, 40.0, 35.0))
    assert mid == pytest.approx(0.5 * (lo + hi), rel=1e-12)
    T = np.array([3.0, 7.0])
    assert np.array_equal(R.linear_density(T, np.full(2, 35.0), s_coef=0.0), R.linear_density(T))


def test_device_eos_host_side():
    from pyocean_amd import eos
    lin = eos.make_eos("linear", rho0=1025.0, t_coef=0.1)
    assert lin.linear and lin.name == "linear" and (lin.rho0, lin.t_coef, lin.t_ref, lin.s_ref) == (1025.0, 0.1, 0.0, 35.0)
    w = eos.make_eos("wright_full")
    assert not w.linear and w.name == "wright_full" and w.rho0 == 1035.0
    with pytest.raises(ValueError, match="only apply to the 'linear'"):
        eos.make_eos("wright_full", t_coef=0.1)
    with pytest.raises(ValueError, match="unknown EOS"):
        eos.make_eos("unesco")
    with pytest.raises(NotImplementedError):
        eos.make_eos("teos10")
    with pytest.raises(ValueError, match="needs salinity"):
        w.density(np.zeros(3))
    from pyocean_amd import vertical
    vc = vertical.VerticalCoordinate(H=None, active=None, rule="hybrid", params={})
    with pytest.raises(ValueError, match="equation of state"):
        vc.regrid_density(np.zeros(3))
    with pytest.raises(ValueError, match="T \\(and S\\) or the layer densities"):
        vc.regrid(None, None)
