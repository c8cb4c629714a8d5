"""
GPU tests of the ALE layer transport (the producer of the tracer step's inputs): every kernel is bit-identical to the
oracle's restatement of the reference arithmetic - transport velocity, transport thickness, PPM edges, the PPM flux
with its Newton correction (same number of passes), the outflow limiter (same passes and count) and the Lagrangian
step - on the reference's own small cases, on regional grids with land, thin z* layers and limited cells, for both
face directions and through the reference's swapped-axes calling convention.  At a full 1024 x 1024 x 50 grid the
size-independent properties are checked: closure of each face column on the free-surface flux, the limiter's
bound, h† >= 0 and the layer-volume identity.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import numpy_ref as R
from pyocean_amd import ale, synth
from tests import cases

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu case]", *a, flush=True)


def _d(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _h(t):
    return t.cpu().numpy()


def _eq(gpu_t, ref, what):
    g = _h(gpu_t)
    assert g.shape == ref.shape, (what, g.shape, ref.shape)
    bad = ~((g == ref) | (np.isnan(g) & np.isnan(ref)))
    assert not bad.any(), f"{what}: {int(bad.sum())} of {g.size} differ, first at {np.argwhere(bad)[0]}, " \
                          f"gpu {g[tuple(np.argwhere(bad)[0])]!r} ref {ref[tuple(np.argwhere(bad)[0])]!r}"


def _inputs(nx, ny, nz, seed):
    a = synth.ale_inputs(nx, ny, nz, seed=seed)
    r = R.ale_layer_transport(a.h, a.H, a.active, a.area, a.dx_t, a.dy_t, a.Lu, a.Lv, a.u_new, a.u_old, a.v_new,
                              a.v_old, a.wet_u, a.wet_v, a.wu, a.wv, a.Du, a.Dv, a.dt, a.theta, a.theta_bc)
    return a, r


@pytest.fixture(scope="module")
def regional():
    return _inputs(64, 48, 10, 1)


# ------------------------------------------------------------------------------------------------ transport velocity
@pytest.mark.parametrize("theta_bc", [None, 0.5])
def test_transport_velocity_matches_the_oracle_bit_for_bit(gpu, regional, theta_bc):
    _say("transport_velocity", theta_bc)
    a, _ = regional
    for u1, u0, wet, w, L, D in ((a.u_new, a.u_old, a.wet_u, a.wu, a.Lu, a.Du),
                                 (a.v_new, a.v_old, a.wet_v, a.wv, a.Lv, a.Dv)):
        ut, Qb = ale.transport_velocity(_d(u1), _d(u0), _d(wet), _d(w), _d(L), _d(D), a.theta, theta_bc)
        rut, rQb = R.transport_velocity(u1, u0, wet, w, L, D, a.theta, theta_bc)
        _eq(ut, rut, "ut")
        _eq(Qb, rQb, "Q_bt")


# ------------------------------------------------------------------------------------------------ transport thickness
@pytest.mark.parametrize("clip", [True, False])
def test_transport_thickness_matches_the_oracle_in_both_directions(gpu, regional, clip):
    _say("face_transport_thickness", clip)
    a, _ = regional
    h = a.h.copy()
    h[3, 5, 7] = np.nan                                                # NaN reads as 0
    H = a.H.copy()
    H[6, 9] = np.nan
    qe, qw = ale.face_transport_thickness(_d(h), _d(H), _d(a.active), clip)
    re_, rw = R.face_transport_thickness(h, H, a.active, clip)
    _eq(qe, re_, "q_e")
    _eq(qw, rw, "q_w")
    # the reference's v-face call: swapped views in, (n, nx, ny+1) out
    qn, qs = ale.face_transport_thickness(_d(h).transpose(1, 2), _d(H).T, _d(a.active).T, clip)
    rn, rs = R.face_transport_thickness(h.swapaxes(1, 2), H.T, a.active.T, clip)
    _eq(qn, rn, "q_n")
    _eq(qs, rs, "q_s")


# ------------------------------------------------------------------------------------------------ PPM
def test_ppm_edges_match_the_oracle_in_both_directions(gpu, regional):
    _say("ppm_edges")
    a, _ = regional
    h = a.h.copy()
    h[2, 10, 10] = np.nan
    h[4, 11, 12] = np.inf
    lo, hi = ale.ppm_edges_x(_d(h), _d(a.active))
    rlo, rhi = R.ppm_edges(h, a.active)
    _eq(lo, rlo, "h_w")
    _eq(hi, rhi, "h_e")
    lo, hi = ale.ppm_edges_x(_d(h).transpose(1, 2), _d(a.active).T)
    rlo, rhi = R.ppm_edges(h.swapaxes(1, 2), a.active.T)
    _eq(lo, rlo, "h_s")
    _eq(hi, rhi, "h_n")


@pytest.mark.parametrize("seed,kw", [(1, {}), (2, {}), (3, {}),
                                     (4, dict(clip=True, land=True, vanish=True)),
                                     (5, dict(clip=True, land=True, vanish=True))])
def test_ppm_transport_matches_the_oracle_on_the_reference_cases(gpu, seed, kw):
    _say("ppm_transport reference case", seed, kw)
    h, ut, L, Qbt, qe, qw, act, dx_t = cases.ppm_random(seed, **kw)
    ref, it_ref = R.ppm_transport_x(h, ut, L, Qbt, qe, qw, act, 60.0, dx_t, return_iterations=True)
    out, it = ale.ppm_transport(_d(h), _d(ut), _d(L), _d(Qbt), _d(qe), _d(qw), _d(act), 60.0, _d(dx_t), 0,
                                return_iterations=True)
    assert it == it_ref
    _eq(out, ref, "Q")


@pytest.mark.parametrize("max_iter", [0, 1, 2, 20])
def test_ppm_transport_stops_where_the_oracle_stops(gpu, max_iter):
    _say("ppm_transport max_iter", max_iter)
    h, ut, L, Qbt, qe, qw, act, dx_t = cases.ppm_random(3, ny=7, nx=13, clip=True, land=True, vanish=True)
    ref, it_ref = R.ppm_transport_x(h, ut, L, Qbt, qe, qw, act, 60.0, dx_t, max_iter, return_iterations=True)
    out, it = ale.ppm_transport(_d(h), _d(ut), _d(L), _d(Qbt), _d(qe), _d(qw), _d(act), 60.0, _d(dx_t), 0,
                                max_iter, return_iterations=True)
    assert it == it_ref
    _eq(out, ref, "Q")


def test_ppm_transport_both_directions_on_a_regional_grid(gpu, regional):
    _say("ppm_transport regional")
    a, r = regional
    Qx, itx = ale.ppm_transport(_d(a.h), _d(r["utx"]), _d(a.Lu), _d(r["Qbx"]), _d(r["qe"]), _d(r["qw"]),
                                _d(a.active), a.dt, _d(a.dx_t), 0, return_iterations=True)
    rx, ritx = R.ppm_transport_x(a.h, r["utx"], a.Lu, r["Qbx"], r["qe"], r["qw"], a.active, a.dt, a.dx_t,
                                 return_iterations=True)
    assert itx == ritx
    _eq(Qx, rx, "Qx (before the limiter)")
    # the reference's calling convention for the v-faces: swapped views, q_n / q_s as (n, nx, ny+1)
    qn_sw, qs_sw = r["qn"].swapaxes(1, 2), r["qs"].swapaxes(1, 2)
    Qy = ale.ppm_transport_x(_d(a.h).transpose(1, 2), _d(r["uty"]).transpose(1, 2), _d(a.Lv).T, _d(r["Qby"]).T,
                             _d(qn_sw), _d(qs_sw), _d(a.active).T, a.dt, _d(a.dy_t).T)
    ry = R.ppm_transport_x(a.h.swapaxes(1, 2), r["uty"].swapaxes(1, 2), a.Lv.T, r["Qby"].T, qn_sw, qs_sw,
                           a.active.T, a.dt, a.dy_t.T)
    _eq(Qy, ry, "Qy (before the limiter)")


# ------------------------------------------------------------------------------------------------ outflow limiter
def _limit_both(h, Qx, Qy, area, dt, act, **kw):
    ref = R.limit_outflow(h, Qx, Qy, area, dt, act, **kw)
    gx, gy = _d(Qx), _d(Qy)
    out = ale.limit_outflow(_d(h), gx, gy, _d(area), dt, _d(act), **kw)
    return ref, out, (gx, gy)


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("periodic", [(False, False), (True, False), (True, True)])
def test_limit_outflow_matches_the_oracle_pass_for_pass(gpu, seed, periodic):
    _say("limit_outflow", seed, periodic)
    h, Qx, Qy, area, dt, act = cases.limiter_case(seed, ny=17, nx=23, periodic=periodic)
    (rx, ry, rn, rst), (gx, gy, gn), _ = _limit_both(h, Qx, Qy, area, dt, act, periodic=periodic)
    assert rst in ("ok", "accepted") and rn > 0
    assert gn == rn
    _eq(gx, rx, "Qx")
    _eq(gy, ry, "Qy")


def test_limit_outflow_reference_cases(gpu):
    _say("limit_outflow reference cases")
    # no-op: the inputs themselves come back
    h, Qx, Qy, area, dt, act = cases.thin_divergent_cell()
    h2 = h.copy()
    h2[1, 0, 1] = 1.0
    (_, _, rn, rst), (gx, gy, gn), (ix, iy) = _limit_both(h2, Qx, Qy, area, dt, act)
    assert rst == "noop" and gn == 0 and gx is ix and gy is iy
    # the thin divergent cell, the counter-flow case, the open edge, the draining same-way layer
    (rx, ry, rn, _), (gx, gy, gn), _ = _limit_both(h, Qx, Qy, area, dt, act)
    assert gn == rn >= 1
    _eq(gx, rx, "Qx thin cell")
    _eq(gy, ry, "Qy thin cell")
    dx, dtt = 1000.0, 10.0
    hh = np.array([[[1e-3, 10.0]], [[10.0, 10.0]]])
    ar = np.full((1, 2), dx * dx)
    for col, q0, q1 in ((1, 2e-3 * dx * dx / dtt, -500.0), (0, -3e-3 * dx * dx / dtt, 500.0)):
        Q = np.zeros((2, 1, 3))
        Q[0, 0, col] = q0
        Q[1, 0, col] = q1
        (rx, _, rn, _), (gx, _, gn), _ = _limit_both(hh, Q, np.zeros((2, 2, 2)), ar, dtt, np.ones((1, 2), bool))
        assert gn == rn
        _eq(gx, rx, f"Qx two-layer case {col}")
    h3 = np.array([[[1e-3, 1e-3, 1e-3]], [[10.0, 10.0, 10.0]], [[10.0, 1e-3, 10.0]]])
    Q3 = np.zeros((3, 1, 4))
    Q3[0, 0, 2] = 3e-3
    Q3[1, 0, 1], Q3[1, 0, 2] = -4.9995, 4.9995
    Q3[2, 0, 1] = 5.0
    (rx, _, rn, _), (gx, _, gn), _ = _limit_both(h3, Q3, np.zeros((3, 2, 3)), np.ones((1, 3)), 1.0,
                                                 np.ones((1, 3), bool))
    assert gn == rn
    _eq(gx, rx, "Qx draining same-way layer")


def test_limit_outflow_raises_when_the_column_cannot_supply_the_flux(gpu):
    _say("limit_outflow raise")
    dx, dt = 1000.0, 10.0
    h = np.array([[[1e-3, 10.0]], [[1e-3, 10.0]]])
    Qx = np.zeros((2, 1, 3))
    Qx[:, 0, 1] = 150.0
    with pytest.raises(RuntimeError, match="barotropic CFL"):
        ale.limit_outflow(_d(h), _d(Qx), _d(np.zeros((2, 2, 2))), _d(np.full((1, 2), dx * dx)), dt,
                          _d(np.ones((1, 2), bool)))


# ------------------------------------------------------------------------------------------------ Lagrangian step
def test_lagrangian_thickness_matches_the_oracle_and_rejects_an_emptied_layer(gpu):
    _say("lagrangian_thickness")
    h = np.full((2, 1, 1), 1.0)
    area = np.full((1, 1), 100.0)
    out = ale.lagrangian_thickness(_d(h), _d(np.array([10.0, -10.0])[:, None, None]), 2.0, _d(area))
    _eq(out, R.lagrangian_thickness(h, np.array([10.0, -10.0])[:, None, None], 2.0, area), "h_dag")
    with pytest.raises(RuntimeError, match="layer 0"):
        ale.lagrangian_thickness(_d(h), _d(np.array([100.0, 0.0])[:, None, None]), 2.0, _d(area))
    tiny = ale.lagrangian_thickness(_d(h), _d(np.array([50.0 + 1e-11, 0.0])[:, None, None]), 2.0, _d(area))
    assert _h(tiny)[0, 0, 0] == 0.0
    rng = np.random.default_rng(3)
    h = rng.uniform(1.0, 2.0, (5, 13, 17))
    dv = np.clip(rng.normal(0.0, 0.3, (5, 13, 17)), -1.0, 1.0)         # |dt dv / A| <= 0.6: nothing empties
    area = rng.uniform(0.5, 2.0, (13, 17))
    area[3, 4] = 0.0
    _eq(ale.lagrangian_thickness(_d(h), _d(dv), 0.3, _d(area)), R.lagrangian_thickness(h, dv, 0.3, area), "h_dag")


# ------------------------------------------------------------------------------------------------ the whole producer
@pytest.mark.parametrize("shape,seed", [((64, 48, 10), 1), ((130, 90, 20), 2)])
def test_layer_transport_step_matches_the_oracle_bit_for_bit(gpu, shape, seed):
    _say("layer transport", shape, seed)
    a, r = _inputs(*shape, seed)
    assert r["n_limited"] > 0                                        # the limiter takes part
    d = {k: _d(v) for k, v in vars(a).items() if isinstance(v, np.ndarray)}
    qe, qw = ale.face_thickness_axis(d["h"], d["H"], d["active"], True, 0)
    qn, qs = ale.face_thickness_axis(d["h"], d["H"], d["active"], True, 1)
    _eq(qn, r["qn"], "q_n")
    lt = ale.LayerTransport(Lu=d["Lu"], Lv=d["Lv"], area=d["area"], active=d["active"], dx_t=d["dx_t"],
                            dy_t=d["dy_t"], dt=a.dt, theta=a.theta)
    out = lt(d["h"], d["u_new"], d["u_old"], d["v_new"], d["v_old"], d["wet_u"], d["wet_v"], d["wu"], d["wv"],
             d["Du"], d["Dv"], qe, qw, qn, qs, theta_bc=a.theta_bc)
    assert out.n_limited == r["n_limited"]
    for k in ("Qx", "Qy", "V0", "V1", "h_dagger"):
        _eq(getattr(out, k), r[k], k)
    _eq(out.ok, r["ok"].astype(np.uint8), "ok")


def test_layer_transport_at_full_size_keeps_closure_and_bounds(gpu):
    _say("layer transport 1024x1024x50")
    a = synth.ale_inputs(1024, 1024, 50, seed=0)
    d = {k: _d(v) for k, v in vars(a).items() if isinstance(v, np.ndarray)}
    qe, qw = ale.face_thickness_axis(d["h"], d["H"], d["active"], True, 0)
    qn, qs = ale.face_thickness_axis(d["h"], d["H"], d["active"], True, 1)
    lt = ale.LayerTransport(Lu=d["Lu"], Lv=d["Lv"], area=d["area"], active=d["active"], dx_t=d["dx_t"],
                            dy_t=d["dy_t"], dt=a.dt, theta=a.theta)
    out = lt(d["h"], d["u_new"], d["u_old"], d["v_new"], d["v_old"], d["wet_u"], d["wet_v"], d["wu"], d["wv"],
             d["Du"], d["Dv"], qe, qw, qn, qs, theta_bc=a.theta_bc, count_iterations=True)
    assert 1 <= min(lt.iterations) and max(lt.iterations) <= 20
    _, Qbx = ale.transport_velocity(d["u_new"], d["u_old"], d["wet_u"], d["wu"], d["Lu"], d["Du"], a.theta,
                                    a.theta_bc)
    _, Qby = ale.transport_velocity(d["v_new"], d["v_old"], d["wet_v"], d["wv"], d["Lv"], d["Dv"], a.theta,
                                    a.theta_bc)
    act = d["active"]
    openx = torch.zeros_like(Qbx, dtype=torch.bool)
    f = torch.arange(1025, device=act.device)
    openx[:] = act[:, (f - 1).clamp(min=0)] & act[:, f.clamp(max=1023)]
    openy = act[(f - 1).clamp(min=0)] & act[f.clamp(max=1023)]
    sx, sy = out.Qx.sum(0), out.Qy.sum(0)
    # each open face column carries the free-surface flux (Newton + residual share, kept by the limiter)
    tol = 1e-12 * torch.maximum(Qbx.abs(), torch.ones_like(Qbx)) * 64
    assert ((sx - Qbx).abs()[openx] <= tol[openx]).all()
    toly = 1e-12 * torch.maximum(Qby.abs(), torch.ones_like(Qby)) * 64
    assert ((sy - Qby).abs()[openy] <= toly[openy]).all()
    assert (out.Qx[:, ~openx] == 0).all() and (out.Qy[:, ~openy] == 0).all()
    assert out.h_dagger.min().item() >= 0.0
    assert torch.equal(out.V0, d["h"] * d["area"][None]) and torch.equal(out.V1, out.h_dagger * d["area"][None])
    assert torch.equal(out.ok.bool(), act.bool()[None] & (out.V1 > 0))


@pytest.mark.parametrize("periodic", [(True, False), (False, True), (True, True)])
def test_periodic_layer_transport_matches_the_oracle_bit_for_bit(gpu, periodic):
    _say("layer transport periodic", periodic)
    a = synth.ale_inputs(64, 48, 10, seed=1)
    r = R.ale_layer_transport(a.h, a.H, a.active, a.area, a.dx_t, a.dy_t, a.Lu, a.Lv, a.u_new, a.u_old, a.v_new,
                              a.v_old, a.wet_u, a.wet_v, a.wu, a.wv, a.Du, a.Dv, a.dt, a.theta, a.theta_bc,
                              periodic=periodic)
    d = {k: _d(v) for k, v in vars(a).items() if isinstance(v, np.ndarray)}
    qe, qw = ale.face_thickness_axis(d["h"], d["H"], d["active"], True, 0, periodic=periodic)
    qn, qs = ale.face_thickness_axis(d["h"], d["H"], d["active"], True, 1, periodic=periodic)
    (re_, rw), (rn, rs) = R.transport_thickness(a.h, a.H, a.active, True, periodic)
    _eq(qe, re_, "q_e")
    _eq(qs, rs, "q_s")
    lt = ale.LayerTransport(Lu=d["Lu"], Lv=d["Lv"], area=d["area"], active=d["active"], dx_t=d["dx_t"],
                            dy_t=d["dy_t"], dt=a.dt, theta=a.theta, periodic=periodic)
    out = lt(d["h"], d["u_new"], d["u_old"], d["v_new"], d["v_old"], d["wet_u"], d["wet_v"], d["wu"], d["wv"],
             d["Du"], d["Dv"], qe, qw, qn, qs, theta_bc=a.theta_bc)
    assert out.n_limited == r["n_limited"]
    for k in ("Qx", "Qy", "V1", "h_dagger"):
        _eq(getattr(out, k), r[k], k)


# ------------------------------------------------------------------------------------------------ upwind thickness flux
def test_upwind_transport_matches_the_oracle_in_both_directions(gpu, regional):
    _say("upwind transport")
    a, r = regional
    utx, Qbx = R.transport_velocity(a.u_new, a.u_old, a.wet_u, a.wu, a.Lu, a.Du, a.theta, a.theta_bc)
    uty, Qby = R.transport_velocity(a.v_new, a.v_old, a.wet_v, a.wv, a.Lv, a.Dv, a.theta, a.theta_bc)
    (qe, qw), (qn, qs) = R.transport_thickness(a.h, a.H, a.active, True)
    refx = R.upwind_transport_x(qe, qw, utx, a.Lu, Qbx)
    refy = R.upwind_transport_x(qn.swapaxes(1, 2), qs.swapaxes(1, 2), uty.swapaxes(1, 2), a.Lv.T, Qby.T)
    _eq(ale.upwind_transport(_d(utx), _d(a.Lu), _d(Qbx), _d(qe), _d(qw), 0), refx, "Qx upwind")
    _eq(ale.upwind_transport(_d(uty), _d(a.Lv), _d(Qby), _d(qn), _d(qs), 1), np.ascontiguousarray(refy.swapaxes(1, 2)),
        "Qy upwind")
    # the reference's calling convention for the v-faces (swapped views in and out)
    got = ale.upwind_transport_x(_d(qn).transpose(1, 2), _d(qs).transpose(1, 2), _d(uty).transpose(1, 2),
                                 _d(a.Lv).T, _d(Qby).T)
    _eq(got, refy, "Qy upwind (swapped)")
    # the reference's lock exchange (test_vertical_ale.py:116-127)
    h = np.array([[[8.0, 0.0]], [[0.0, 8.0]]])
    le, lw = R.face_transport_thickness(h, h.sum(0), np.ones((1, 2), bool), True)
    ut = np.zeros((2, 1, 3))
    ut[:, 0, 1] = [0.2, -0.2]
    _eq(ale.upwind_transport_x(_d(le), _d(lw), _d(ut), _d(np.full((1, 3), 100.0)), _d(np.zeros((1, 3)))),
        R.upwind_transport_x(le, lw, ut, np.full((1, 3), 100.0), np.zeros((1, 3))), "lock exchange")


@pytest.mark.parametrize("periodic", [(False, False), (True, True)])
def test_upwind_layer_transport_matches_the_oracle_bit_for_bit(gpu, periodic):
    _say("layer transport upwind", periodic)
    a = synth.ale_inputs(130, 90, 20, seed=2)
    r = R.ale_layer_transport(a.h, a.H, a.active, a.area, a.dx_t, a.dy_t, a.Lu, a.Lv, a.u_new, a.u_old, a.v_new,
                              a.v_old, a.wet_u, a.wet_v, a.wu, a.wv, a.Du, a.Dv, a.dt, a.theta, a.theta_bc,
                              periodic=periodic, thickness_flux="upwind")
    d = {k: _d(v) for k, v in vars(a).items() if isinstance(v, np.ndarray)}
    qe, qw = ale.face_thickness_axis(d["h"], d["H"], d["active"], True, 0, periodic=periodic)
    qn, qs = ale.face_thickness_axis(d["h"], d["H"], d["active"], True, 1, periodic=periodic)
    lt = ale.LayerTransport(Lu=d["Lu"], Lv=d["Lv"], area=d["area"], active=d["active"], dx_t=d["dx_t"],
                            dy_t=d["dy_t"], dt=a.dt, theta=a.theta, periodic=periodic, thickness_flux="upwind")
    out = lt(d["h"], d["u_new"], d["u_old"], d["v_new"], d["v_old"], d["wet_u"], d["wet_v"], d["wu"], d["wv"],
             d["Du"], d["Dv"], qe, qw, qn, qs, theta_bc=a.theta_bc, count_iterations=True)
    assert out.n_limited == r["n_limited"] and lt.iterations == (0, 0)
    for k in ("Qx", "Qy", "V0", "V1", "h_dagger"):
        _eq(getattr(out, k), r[k], k)
    _eq(out.ok, r["ok"].astype(np.uint8), "ok")
    with pytest.raises(ValueError, match="thickness_flux"):
        ale.LayerTransport(Lu=d["Lu"], Lv=d["Lv"], area=d["area"], active=d["active"], dx_t=d["dx_t"],
                           dy_t=d["dy_t"], dt=a.dt, theta=a.theta, thickness_flux="weno")


@pytest.mark.parametrize("shape", [(1, 1, 1), (1, 3, 2), (4, 1, 7), (2, 5, 1)])
def test_upwind_transport_on_tiny_and_ragged_shapes(gpu, shape):
    _say("upwind shapes", shape)
    n, ny, nx = shape
    rng = np.random.default_rng(sum(shape))
    h = rng.uniform(0.0, 3.0, shape)
    act = rng.random((ny, nx)) > 0.2
    qe, qw = R.face_transport_thickness(h, h.sum(0) + 1.0, act, True)
    ut = rng.normal(0.0, 0.2, (n, ny, nx + 1))
    L, Qb = rng.uniform(1.0, 2.0, (ny, nx + 1)), rng.normal(0.0, 3.0, (ny, nx + 1))
    _eq(ale.upwind_transport(_d(ut), _d(L), _d(Qb), _d(qe), _d(qw), 0), R.upwind_transport_x(qe, qw, ut, L, Qb), "Qx")
    (qn, qs) = (np.ascontiguousarray(q.swapaxes(1, 2)) for q in
                R.face_transport_thickness(h.swapaxes(1, 2), (h.sum(0) + 1.0).T, act.T, True))
    vt = rng.normal(0.0, 0.2, (n, ny + 1, nx))
    Lv, Qv = rng.uniform(1.0, 2.0, (ny + 1, nx)), rng.normal(0.0, 3.0, (ny + 1, nx))
    ref = R.upwind_transport_x(qn.swapaxes(1, 2), qs.swapaxes(1, 2), vt.swapaxes(1, 2), Lv.T, Qv.T).swapaxes(1, 2)
    _eq(ale.upwind_transport(_d(vt), _d(Lv), _d(Qv), _d(qn), _d(qs), 1), np.ascontiguousarray(ref), "Qy")


# ------------------------------------------------------------------------------------------------ open-edge faces
def _open_fields(a, rng):
    ny, nx = a.H.shape
    Du = np.where(rng.random((ny, nx + 1)) < 0.9, rng.uniform(1.0, 50.0, (ny, nx + 1)), 0.0)
    Dv = np.where(rng.random((ny + 1, nx)) < 0.9, rng.uniform(1.0, 50.0, (ny + 1, nx)), 0.0)
    Du[0, 0] = np.nan
    Qu, Qv = rng.normal(0.0, 5e4, (ny, nx + 1)), rng.normal(0.0, 5e4, (ny + 1, nx))
    return Du, Qu, Dv, Qv


@pytest.mark.parametrize("edges", [("west",), ("east", "south"), ("west", "east", "south", "north")])
def test_open_face_transports_match_the_oracle(gpu, regional, edges):
    _say("open face transports", edges)
    a, _ = regional
    rng = np.random.default_rng(len(edges))
    Du, Qu, Dv, Qv = _open_fields(a, rng)
    n, ny, nx = a.h.shape
    Qx0, Qy0 = rng.normal(size=(n, ny, nx + 1)), rng.normal(size=(n, ny + 1, nx))
    rx, ry = R.open_face_transports(Qx0.copy(), Qy0.copy(), a.h, a.u_old, a.v_old, a.u_new, a.v_new, edges, Du, Qu,
                                    Dv, Qv, a.Lu, a.Lv, a.theta, a.theta_bc)
    gx, gy = _d(Qx0), _d(Qy0)
    ale.open_face_transports(gx, gy, _d(a.h), _d(a.u_old), _d(a.v_old), _d(a.u_new), _d(a.v_new), edges, _d(Du),
                             _d(Qu), _d(Dv), _d(Qv), _d(a.Lu), _d(a.Lv), a.theta, a.theta_bc)
    _eq(gx, rx, "Qx")
    _eq(gy, ry, "Qy")


def test_layer_transport_with_open_edges_matches_the_oracle_bit_for_bit(gpu):
    _say("layer transport with open edges")
    a = synth.ale_inputs(64, 48, 10, seed=1)
    rng = np.random.default_rng(9)
    Du, Qu, Dv, Qv = _open_fields(a, rng)
    Qu, Qv = Qu * 1e-2, Qv * 1e-2
    edges = ("west", "north")
    r = R.ale_layer_transport(a.h, a.H, a.active, a.area, a.dx_t, a.dy_t, a.Lu, a.Lv, a.u_new, a.u_old, a.v_new,
                              a.v_old, a.wet_u, a.wet_v, a.wu, a.wv, a.Du, a.Dv, a.dt, a.theta, a.theta_bc,
                              open_=edges, open_Du=Du, open_Qu=Qu, open_Dv=Dv, open_Qv=Qv)
    d = {k: _d(v) for k, v in vars(a).items() if isinstance(v, np.ndarray)}
    qe, qw = ale.face_thickness_axis(d["h"], d["H"], d["active"], True, 0)
    qn, qs = ale.face_thickness_axis(d["h"], d["H"], d["active"], True, 1)
    lt = ale.LayerTransport(Lu=d["Lu"], Lv=d["Lv"], area=d["area"], active=d["active"], dx_t=d["dx_t"],
                            dy_t=d["dy_t"], dt=a.dt, theta=a.theta)
    out = lt(d["h"], d["u_new"], d["u_old"], d["v_new"], d["v_old"], d["wet_u"], d["wet_v"], d["wu"], d["wv"],
             d["Du"], d["Dv"], qe, qw, qn, qs, theta_bc=a.theta_bc, open_=edges, open_Du=_d(Du), open_Qu=_d(Qu),
             open_Dv=_d(Dv), open_Qv=_d(Qv))
    assert out.n_limited == r["n_limited"]
    for k in ("Qx", "Qy", "V0", "V1", "h_dagger"):
        _eq(getattr(out, k), r[k], k)
    assert (r["Qx"][:, :, 0] != 0).any() and (r["Qy"][:, -1, :] != 0).any()
