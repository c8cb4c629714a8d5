"""
GPU parity of the 'upwind3' / 'upwind1' tracer schemes of TracerAdvection.advect, rhs and flux_x (UP3 / donor faces
of _up3_faces and rhs_numpy, sigma_baroclinic.py:164-187, tracers.py:139-151, 342-375), of the centred faces
(rhs / flux_x with centred=True, _centred_faces sigma_baroclinic.py:632-652) and of the 'pc' step (advect_numpy,
tracers.py:303-330) against the oracle, bit for bit (the bit patterns, so signed zeros count), through the C-ABI;
and the reference's dispatch of the FCT step and faces_h for the upwind schemes (WENO5-JS).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import numpy_ref as R
from pyocean_amd import Periodic, ops
from pyocean_amd.tracers import OpenEdges, TracerAdvection
from tests import cases

pytestmark = pytest.mark.gpu

EDGES = ("west", "east", "south", "north")
ORDER = {"upwind3": 3, "upwind1": 1}


def _say(*a):
    print("[gpu case]", *a, flush=True)


def _adv(scheme, open_mask=0, dt=30.0, per=None, time="ssprk3"):
    adv = TracerAdvection(scheme=scheme, time=time, dt=dt,
                          ob=OpenEdges([e for b, e in zip((1, 2, 4, 8), EDGES) if open_mask & b]))
    if per:
        adv.per = Periodic(per)
    return adv


def _same_bits(a, b):
    a, b = np.ascontiguousarray(a, float), np.ascontiguousarray(b, float)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def _fz(shape, seed, flat=False):
    nz, ny, nx = shape
    Fz = np.random.default_rng(seed).normal(0.0, 2e3, (nz + 1, ny, nx))
    Fz[0] = Fz[-1] = 0.0
    if flat:
        Fz[:] = 0.0
    return Fz


@pytest.mark.parametrize("scheme", ["upwind3", "upwind1"])
@pytest.mark.parametrize("variant", ["plain", "inflow", "flat", "closed"])
def test_rk3_step_matches_the_oracle_bit_for_bit(gpu, scheme, variant):
    # "closed": edges closed while the edge transports stay nonzero, so the closed faces are Q * 0 (signed zeros)
    case = cases.fused_fct_case(open_=True, seed=11)
    mask = 0 if variant == "closed" else case["open_mask"]
    _say("up rk3", scheme, variant)
    nz, ny, nx = case["V0"].shape
    Fz = _fz((nz, ny, nx), 4, flat=variant == "flat")
    inflow = case["C"] + 1.0 if variant == "inflow" else None
    adv = _adv(scheme, mask, dt=case["dt"])
    args = (case["C"], case["V0"], case["V1"], case["Qx"], case["Qy"], Fz, case["ok"])
    got, fxi = adv.advect(*args, inflow=inflow, flux_face_i=nx)
    ref, fref = R.advect_rk3(*args, case["dt"], inflow=inflow, open_mask=mask, flux_face_i=nx,
                             up_order=ORDER[scheme])
    assert _same_bits(got, ref)
    assert _same_bits(fxi, fref)
    got_i = adv.advect(*args, inflow=inflow, flux_face_i=3)[1]
    assert _same_bits(got_i, R.advect_rk3(*args, case["dt"], inflow=inflow, open_mask=mask, flux_face_i=3,
                                          up_order=ORDER[scheme])[1])


@pytest.mark.parametrize("scheme", ["upwind3", "upwind1"])
def test_rhs_and_flux_x_match_the_oracle(gpu, scheme):
    case = cases.fused_fct_case(open_=True, seed=12)
    _say("up rhs", scheme)
    nz, ny, nx = case["V0"].shape
    adv = _adv(scheme, case["open_mask"])
    ext = case["C"] - 0.5
    for flat in (False, True):
        Fz = _fz((nz, ny, nx), 6, flat=flat)
        got = adv.rhs(case["C"], case["Qx"], case["Qy"], Fz, case["ok"], inflow=ext)
        ref = R.rhs_up(case["C"], case["Qx"], case["Qy"], Fz, case["ok"], ext, case["open_mask"], ORDER[scheme])
        assert _same_bits(got, ref)
    fx = adv.flux_x(case["C"], case["Qx"], case["Qy"], case["ok"], inflow=ext)
    assert _same_bits(fx, R.faces_up_h(case["C"], case["Qx"], case["Qy"], case["ok"], ext, case["open_mask"],
                                       ORDER[scheme])[0])


@pytest.mark.parametrize("shape", [(1, 1, 1, 1), (1, 2, 1, 7), (1, 3, 6, 1), (3, 5, 37, 70), (2, 3, 130, 129)])
def test_ragged_and_tiny_shapes(gpu, shape):
    _say("up shape", shape)
    nt, nz, ny, nx = shape
    rng = np.random.default_rng(sum(shape))
    C = rng.normal(5.0, 1.0, shape)
    V = rng.uniform(0.8, 1.2, (nz, ny, nx)) * 1e6
    Qx = rng.uniform(-0.3, 0.3, (nz, ny, nx + 1)) * 1e4
    Qy = rng.uniform(-0.3, 0.3, (nz, ny + 1, nx)) * 1e4
    Fz = rng.uniform(-0.3, 0.3, (nz + 1, ny, nx)) * 1e4
    V1 = V - 30.0 * (Qx[..., 1:] - Qx[..., :-1] + Qy[:, 1:] - Qy[:, :-1] + Fz[1:] - Fz[:-1])
    ok = rng.random((nz, ny, nx)) > 0.15
    for scheme in ("upwind3", "upwind1"):
        got = _adv(scheme, 15).advect(C, V, V1, Qx, Qy, Fz, ok)
        assert _same_bits(got, R.advect_rk3(C, V, V1, Qx, Qy, Fz, ok, 30.0, open_mask=15, up_order=ORDER[scheme]))


@pytest.mark.parametrize("per", [("x",), ("x", "y")])
def test_periodic_pad_crop_matches_the_oracle(gpu, per):
    _say("up periodic", per)
    px, py = "x" in per, "y" in per
    rng = np.random.default_rng(9)
    nt, nz, ny, nx = 2, 3, 20, 24
    V = rng.uniform(0.8, 1.2, (nz, ny, nx)) * 1e6
    Qx = rng.uniform(-0.3, 0.3, (nz, ny, nx + 1)) * 1e4
    Qy = rng.uniform(-0.3, 0.3, (nz, ny + 1, nx)) * 1e4
    Qx[..., -1] = Qx[..., 0]
    if py:
        Qy[:, -1] = Qy[:, 0]
    else:
        Qy[:, [0, -1]] = 0.0
    V1 = V - 30.0 * (Qx[..., 1:] - Qx[..., :-1] + Qy[:, 1:] - Qy[:, :-1])
    ok = V1 > 0
    C = rng.normal(5.0, 1.0, (nt, nz, ny, nx))
    Fz = _fz((nz, ny, nx), 3)
    H = R.TRACER_HALO
    P = lambda a, k: R.pad(a, k, H, px, py)                              # noqa: E731
    got = _adv("upwind3", per=per).advect(C, V, V1, Qx, Qy, Fz, ok)
    ref = R.advect_rk3(P(C, "c"), P(V, "c"), P(V1, "c"), P(Qx, "u"), P(Qy, "v"), P(Fz, "c"), P(ok, "c"), 30.0,
                       up_order=3)
    assert _same_bits(got, R.crop(ref, H, px, py))


def test_fct_step_and_faces_h_run_weno5js_for_the_upwind_schemes(gpu, oracle_c):
    case = cases.fused_fct_case(open_=True, seed=13)
    _say("up fct dispatch")
    args = (case["C"], case["V0"], case["V1"], case["Qx"], case["Qy"], case["ok"])
    got = _adv("upwind3", case["open_mask"], dt=case["dt"]).advect_fct(*args)
    want = _adv("weno5js", case["open_mask"], dt=case["dt"]).advect_fct(*args)
    assert np.array_equal(got, want)
    ref = oracle_c.advect_fct(*args, case["dt"], open_mask=case["open_mask"], z_variant=False)
    assert np.array_equal(got, ref)
    fx, fy = _adv("upwind1", case["open_mask"]).faces_h(case["C"], case["Qx"], case["Qy"], case["ok"])
    rx, ry = R.faces_h(case["C"], case["Qx"], case["Qy"], case["ok"], case["C"], case["open_mask"], z_variant=False)
    assert np.array_equal(fx, rx) and np.array_equal(fy, ry)


def test_c_abi_rejects_two_face_schemes_and_a_missing_fz(gpu):
    _say("up flag errors")
    dev = torch.device("cuda:0")
    nt, n, ny, nx = 1, 2, 3, 4
    z = lambda *s: torch.zeros(s, dtype=torch.float64, device=dev)      # noqa: E731
    C, V, Qx, Qy, Fz = z(nt, n, ny, nx), z(n, ny, nx) + 1.0, z(n, ny, nx + 1), z(n, ny + 1, nx), z(n + 1, ny, nx)
    ok = torch.ones((n, ny, nx), dtype=torch.uint8, device=dev)
    with pytest.raises(ValueError, match="more than one face scheme"):
        ops.advect_rk3(C, V, V, Qx, Qy, Fz, ok, 30.0, flags=ops.flags_of(js=True, up=3))
    with pytest.raises(ValueError, match="not valid"):
        ops.advect_rk3(C, V, V, Qx, Qy, Fz, ok, 30.0, flags=ops.flags_of(unlimited=True, up=1))
    with pytest.raises(ValueError, match="need Fz"):
        ops.rhs(C, Qx, Qy, None, ok, flags=ops.flags_of(up=3))
    with pytest.raises(ValueError, match="flag"):
        ops.advect_fct(C, V, V, Qx, Qy, ok, 30.0, flags=ops.flags_of(up=3))
    torch.cuda.synchronize()


# ------------------------------------------------------------------------------------------------ centred, 'pc'
UPO = {"weno5": 0, "weno5js": 0, "upwind3": 3, "upwind1": 1}


@pytest.mark.parametrize("scheme", ["weno5", "weno5js", "upwind3", "upwind1"])
@pytest.mark.parametrize("variant", ["plain", "inflow", "flat", "closed"])
def test_pc_step_matches_the_oracle_bit_for_bit(gpu, scheme, variant):
    case = cases.fused_fct_case(open_=True, seed=21)
    mask = 0 if variant == "closed" else case["open_mask"]
    _say("pc", scheme, variant)
    nz, ny, nx = case["V0"].shape
    Fz = _fz((nz, ny, nx), 5, flat=variant == "flat")
    inflow = case["C"] + 1.0 if variant == "inflow" else None
    adv = _adv(scheme, mask, dt=case["dt"], time="pc")
    args = (case["C"], case["V0"], case["V1"], case["Qx"], case["Qy"], Fz, case["ok"])
    got, fxi = adv.advect(*args, inflow=inflow, flux_face_i=nx)
    ref, fref = R.advect_pc(*args, case["dt"], inflow=inflow, open_mask=mask, z_variant=scheme == "weno5",
                            flux_face_i=nx, up_order=UPO[scheme])
    assert _same_bits(got, ref)
    assert _same_bits(fxi, fref)
    assert _same_bits(adv.advect(*args, inflow=inflow), ref)


def test_centred_rhs_and_flux_x_match_the_oracle(gpu):
    case = cases.fused_fct_case(open_=True, seed=22)
    _say("centred rhs")
    nz, ny, nx = case["V0"].shape
    ext = case["C"] - 0.5
    for scheme in ("weno5", "upwind1"):
        adv = _adv(scheme, case["open_mask"])
        for flat in (False, True):
            Fz = _fz((nz, ny, nx), 7, flat=flat)
            got = adv.rhs(case["C"], case["Qx"], case["Qy"], Fz, case["ok"], centred=True, inflow=ext)
            ref = R.rhs_up(case["C"], case["Qx"], case["Qy"], Fz, case["ok"], ext, case["open_mask"], "c")
            assert _same_bits(got, ref)
        fx = adv.flux_x(case["C"], case["Qx"], case["Qy"], case["ok"], centred=True, inflow=ext)
        assert _same_bits(fx, R.faces_up_h(case["C"], case["Qx"], case["Qy"], case["ok"], ext, case["open_mask"],
                                           "c")[0])


@pytest.mark.parametrize("shape", [(1, 1, 1, 1), (1, 2, 1, 7), (1, 7, 6, 1), (3, 5, 37, 70), (2, 3, 130, 129)])
def test_pc_ragged_and_tiny_shapes(gpu, shape):
    _say("pc shape", shape)
    nt, nz, ny, nx = shape
    rng = np.random.default_rng(sum(shape) + 1)
    C = rng.normal(5.0, 1.0, shape)
    V = rng.uniform(0.8, 1.2, (nz, ny, nx)) * 1e6
    Qx = rng.uniform(-0.3, 0.3, (nz, ny, nx + 1)) * 1e4
    Qy = rng.uniform(-0.3, 0.3, (nz, ny + 1, nx)) * 1e4
    Fz = rng.uniform(-0.3, 0.3, (nz + 1, ny, nx)) * 1e4
    V1 = V - 30.0 * (Qx[..., 1:] - Qx[..., :-1] + Qy[:, 1:] - Qy[:, :-1] + Fz[1:] - Fz[:-1])
    ok = rng.random((nz, ny, nx)) > 0.15
    for scheme in ("weno5", "upwind3"):
        got = _adv(scheme, 15, time="pc").advect(C, V, V1, Qx, Qy, Fz, ok)
        ref = R.advect_pc(C, V, V1, Qx, Qy, Fz, ok, 30.0, open_mask=15, up_order=UPO[scheme])
        assert _same_bits(got, ref)


def test_pc_periodic_pad_crop_matches_the_oracle(gpu):
    _say("pc periodic")
    rng = np.random.default_rng(23)
    nt, nz, ny, nx = 2, 3, 20, 24
    V = rng.uniform(0.8, 1.2, (nz, ny, nx)) * 1e6
    Qx = rng.uniform(-0.3, 0.3, (nz, ny, nx + 1)) * 1e4
    Qy = rng.uniform(-0.3, 0.3, (nz, ny + 1, nx)) * 1e4
    Qx[..., -1] = Qx[..., 0]
    Qy[:, -1] = Qy[:, 0]
    V1 = V - 30.0 * (Qx[..., 1:] - Qx[..., :-1] + Qy[:, 1:] - Qy[:, :-1])
    ok = V1 > 0
    C = rng.normal(5.0, 1.0, (nt, nz, ny, nx))
    Fz = _fz((nz, ny, nx), 8)
    H = R.TRACER_HALO
    P = lambda a, k: R.pad(a, k, H, True, True)                          # noqa: E731
    got = _adv("weno5", per=("x", "y"), time="pc").advect(C, V, V1, Qx, Qy, Fz, ok)
    ref = R.advect_pc(P(C, "c"), P(V, "c"), P(V1, "c"), P(Qx, "u"), P(Qy, "v"), P(Fz, "c"), P(ok, "c"), 30.0)
    assert _same_bits(got, R.crop(ref, H, True, True))
    # the face column of a periodic step: index shifted by the halo, the y halo cropped (tracers.py:80-102)
    for scheme, time_, run in (("upwind1", "pc", R.advect_pc), ("upwind3", "ssprk3", R.advect_rk3)):
        got, fxi = _adv(scheme, per=("x", "y"), time=time_).advect(C, V, V1, Qx, Qy, Fz, ok, flux_face_i=3)
        ref, fref = run(P(C, "c"), P(V, "c"), P(V1, "c"), P(Qx, "u"), P(Qy, "v"), P(Fz, "c"), P(ok, "c"), 30.0,
                        flux_face_i=3 + H, up_order=ORDER[scheme])
        assert _same_bits(got, R.crop(ref, H, True, True))
        assert _same_bits(fxi, fref[..., H:-H])
    adv = _adv("upwind3", time="pc")
    assert _same_bits(adv.advect_numpy(C, V, V1, Qx, Qy, Fz, ok), adv.advect(C, V, V1, Qx, Qy, Fz, ok))
