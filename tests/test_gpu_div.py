"""
The fast FP64 divisions of the kernels against the IEEE quotient, bit for bit, on operands that sweep every guard
edge (GPU).  The FP64 kernels divide finite values by the WENO and F_H constants with ``div_const`` (the IEEE
division below 2^-960 in magnitude, except for zero), by data with ``div_nr`` inside the WENO face's fast range (stencil magnitudes in 2^-400 .. 2^100) and with
``qdiv`` in the remap (both operands in 2^-500 .. 2^500); everywhere else they take the IEEE division.  The host's
NumPy quotient is the IEEE reference.  A second test runs the WENO faces and the FCT step on tracer fields scaled
across the face's fast-range edges, against the oracle.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from pyocean_amd import _native, ops

pytestmark = pytest.mark.gpu

EDGES = [-1074, -1073, -1060, -1023, -1022, -1021, -1000, -970, -962, -961, -960, -959, -958, -700, -600, -502,
         -501, -500, -499, -498, -404, -402, -401, -400, -399, -398, -200, -133, -60, -1, 0, 1, 60, 98, 99, 100,
         101, 102, 200, 498, 499, 500, 501, 502, 900, 1000, 1021, 1022, 1023]


def _operands(seed=0, per_edge=24):
    """x of both signs and y > 0 with exponents on and next to every guard edge, random mantissas (and the
    mantissa extremes), plus the special values."""
    rng = np.random.default_rng(seed)
    xs, ys = [], []
    for e in EDGES:
        m = np.concatenate([[1.0, np.nextafter(2.0, 0.0), 1.5], rng.uniform(1.0, 2.0, per_edge - 3)])
        v = np.ldexp(m, e)                                   # ldexp handles the subnormal exponents
        xs.append(v)
        xs.append(-v)
        ys.append(v)
    x = np.concatenate(xs + [np.array([0.0, -0.0, np.inf, -np.inf, np.nan, np.finfo(float).max,
                                       -np.finfo(float).max, 6.0, -3.0, 12.0, 5e-324, -5e-324])])
    y = np.concatenate(ys + [np.array([6.0, 3.0, 1.0, 1e-40, 1e-12, np.finfo(float).max, 5e-324, 2.2250738585072014e-308])])
    X, Y = np.meshgrid(x, y, indexing="ij")
    return X.ravel(), Y.ravel()


def _same(a, b):
    """bit patterns equal, NaN matching any NaN"""
    nan = np.isnan(a) & np.isnan(b)
    return nan | (a.view(np.int64) == b.view(np.int64))


def test_fast_divisions_equal_the_ieee_quotient_bit_for_bit(gpu):
    x, y = _operands()
    n = x.size
    print(f"[gpu case] selftest div n={n}", flush=True)
    dx, dy = torch.from_numpy(x).to(gpu), torch.from_numpy(y).to(gpu)
    out = torch.empty(6 * n, dtype=torch.float64, device=gpu)
    L = _native.lib()
    _native.check(L.ocx_selftest_div_f64(dx.data_ptr(), dy.data_ptr(), n, out.data_ptr(),
                                         torch.cuda.current_stream().cuda_stream), "ocx_selftest_div_f64")
    torch.cuda.synchronize()
    o = out.cpu().numpy().reshape(6, n)
    with np.errstate(all="ignore"):
        q6, q3, q = x / 6.0, x / 3.0, x / y
    assert _same(o[4], q).all(), "the device's IEEE division differs from the host's"
    # div_const is defined for finite x (the kernels divide finite sums of finite values by 6 and 3): there it is the
    # IEEE quotient, zeros of both signs and subnormals included; a NaN stays a NaN
    fin = np.isfinite(x)
    bad = fin & ~_same(o[0], q6)
    assert not bad.any(), f"div_const(x, 6) differs at x = {x[bad][:5]}"
    bad = fin & ~_same(o[1], q3)
    assert not bad.any(), f"div_const(x, 3) differs at x = {x[bad][:5]}"
    assert np.isnan(o[0][np.isnan(x)]).all() and np.isnan(o[1][np.isnan(x)]).all()
    ax = np.abs(x)
    rng = ((ax >= 2.0 ** -500) & (ax <= 2.0 ** 500) | (x == 0.0)) & (y >= 2.0 ** -500) & (y <= 2.0 ** 500)
    assert rng.sum() > 10000
    bad = rng & ~_same(o[2], q)
    assert not bad.any(), f"div_nr differs at (x, y) = {list(zip(x[bad][:5], y[bad][:5]))}"
    bad = ~_same(o[3], q)
    assert not bad.any(), f"qdiv differs at (x, y) = {list(zip(x[bad][:5], y[bad][:5]))}"
    # the same with the divisor's reciprocal refined first (the vertical diffusion solve's pivots)
    bad = ~_same(o[5], q)
    assert not bad.any(), f"qdiv_r differs at (x, y) = {list(zip(x[bad][:5], y[bad][:5]))}"


@pytest.mark.parametrize("scale_exp", [-404, -401, -400, -399, -396, 96, 99, 100, 101, 104, -1000, 700])
def test_weno_faces_and_fct_step_on_fields_scaled_across_the_fast_range_edges(gpu, oracle_c, scale_exp):
    # the WENO face takes its fast divisions when the stencil's largest magnitude lies in 2^-400 .. 2^100: a field
    # spanning a factor ~16 scaled next to either edge puts stencils on both sides of it
    from tests import cases
    case = cases.fused_fct_case(open_=True, nx=70, ny=40, nz=4)
    print(f"[gpu case] scaled fields 2^{scale_exp}", flush=True)
    C = np.ldexp(case["C"], scale_exp)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu)       # noqa: E731
    okd = d(case["ok"].astype(np.uint8))
    fx, fy = ops.faces_h(d(C), d(case["Qx"]), d(case["Qy"]), okd, flags=case["open_mask"])
    rfx, rfy = oracle_c.faces_h(C, case["Qx"], case["Qy"], case["ok"], C, case["open_mask"])
    assert _same(fx.cpu().numpy(), rfx).all() and _same(fy.cpu().numpy(), rfy).all()
    if -1000 < scale_exp < 700:
        out = ops.advect_fct(d(C), d(case["V0"]), d(case["V1"]), d(case["Qx"]), d(case["Qy"]), okd, case["dt"],
                             flags=case["open_mask"])
        ref = oracle_c.advect_fct(C, case["V0"], case["V1"], case["Qx"], case["Qy"], case["ok"], case["dt"],
                                  open_mask=case["open_mask"])
        assert _same(out.cpu().numpy(), ref).all()


@pytest.mark.parametrize("scale_exp", [-402, -400, -398, 98, 100, 102])
@pytest.mark.parametrize("variant", ["z", "js"])
def test_rk3_step_with_z_sweep_on_fields_scaled_across_the_fast_range_edges(gpu, oracle_c, scale_exp, variant):
    # the RK3 stage kernel checks the fast range once per layer tile (x and y faces) and per column window (z faces):
    # with a field spanning a factor ~16 next to an edge some tiles and windows are inside the range and some are not
    from tests import cases
    case = cases.fused_fct_case(open_=True, nx=70, ny=40, nz=6, seed=5)
    print(f"[gpu case] rk3 scaled fields 2^{scale_exp} {variant}", flush=True)
    nz, ny, nx = case["V0"].shape
    C = np.ldexp(case["C"], scale_exp)
    Fz = np.random.default_rng(4).normal(0.0, 2e3, (nz + 1, ny, nx))
    Fz[0] = Fz[-1] = 0.0
    flags = case["open_mask"] | (16 if variant == "js" else 0)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu)       # noqa: E731
    out = ops.advect_rk3(d(C), d(case["V0"]), d(case["V1"]), d(case["Qx"]), d(case["Qy"]), d(Fz),
                         d(case["ok"].astype(np.uint8)), case["dt"], flags=flags)
    ref = oracle_c.advect_rk3(C, case["V0"], case["V1"], case["Qx"], case["Qy"], Fz, case["ok"], case["dt"],
                              open_mask=case["open_mask"], z_variant=variant == "z")
    assert _same(out.cpu().numpy(), ref).all()
