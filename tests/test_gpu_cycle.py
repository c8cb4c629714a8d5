"""
The whole ALE tracer cycle on the device against the same cycle through the oracle, bit for bit: the layer
transport (PPM fluxes with the Newton correction, the outflow limiter, the Lagrangian step), the FCT advection of a
tracer stack, the along-layer and implicit vertical diffusion, and the z* regrid with the PPM_H4 remap.  Each phase
hands the next one its results on the device; the free surface that the regrid needs is the one value formed on
the host for both sides.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import cref
from oracle import numpy_ref as R
from pyocean_amd import ale, ops, synth, vertical

pytestmark = pytest.mark.gpu


def _d(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _eq(gpu_t, ref, what):
    g = gpu_t.cpu().numpy()
    assert g.shape == ref.shape, (what, g.shape, ref.shape)
    bad = ~((g == ref) | (np.isnan(g) & np.isnan(ref)))
    if bad.any():
        w = tuple(np.argwhere(bad)[0])
        raise AssertionError(f"{what}: {int(bad.sum())} of {g.size} differ, first at {w}: gpu {g[w]!r} ref {ref[w]!r}")


@pytest.mark.parametrize("shape,nt", [((72, 56, 12), 3), ((130, 90, 20), 2)])
def test_ale_tracer_cycle_matches_the_oracle_bit_for_bit(gpu, shape, nt):
    print("[gpu case] ALE cycle", shape, nt, flush=True)
    nx, ny, nz = shape
    a = synth.ale_inputs(nx, ny, nz, seed=3)
    rng = np.random.default_rng(8)
    act = a.active.astype(bool)
    C0 = 10.0 + rng.normal(0.0, 1.0, (nt, nz, ny, nx)) * act
    cx = a.Lu / np.concatenate([a.dx_t[:, :1], 0.5 * (a.dx_t[:, 1:] + a.dx_t[:, :-1]), a.dx_t[:, -1:]], 1)
    cy = a.Lv / np.concatenate([a.dy_t[:1], 0.5 * (a.dy_t[1:] + a.dy_t[:-1]), a.dy_t[-1:]], 0)
    okx = np.zeros((ny, nx + 1), bool)
    okx[:, 1:-1] = act[:, 1:] & act[:, :-1]
    oky = np.zeros((ny + 1, nx), bool)
    oky[1:-1] = act[1:] & act[:-1]
    M = ny * nx
    nu = np.full((nz + 1, M), 1e-3)
    dz_top = np.full(nz, 2000.0 / nz)
    # the oracle's cycle
    fl = R.ale_layer_transport(a.h, a.H, a.active, a.area, a.dx_t, a.dy_t, a.Lu, a.Lv, a.u_new, a.u_old, a.v_new,
                               a.v_old, a.wet_u, a.wet_v, a.wu, a.wv, a.Du, a.Dv, a.dt, a.theta, a.theta_bc)
    C1 = cref.advect_fct(C0, fl["V0"], fl["V1"], fl["Qx"], fl["Qy"], fl["ok"], a.dt)
    hd = fl["h_dagger"]
    C2 = R.diffuse_h_ale(C1, C1, hd, act, cx, cy, okx, oky, fl["V1"], 100.0, a.dt)
    C2 = R.vertical_implicit(C2.reshape(nt, nz, M), np.maximum(hd, 1e-6).reshape(nz, M), nu, np.zeros(M),
                             a.dt).reshape(nt, nz, ny, nx)
    eta = hd.sum(0) - a.H
    h_ref, wet = cref.regrid(a.H, eta, a.active, hd, "zstar", dz_top)
    C_ref = cref.remap_columns(hd, C2, h_ref, cref.PPM_H4, wet)
    # the same cycle on the device
    d = {k: _d(v) for k, v in vars(a).items() if isinstance(v, np.ndarray)}
    lt = ale.LayerTransport(Lu=d["Lu"], Lv=d["Lv"], area=d["area"], active=d["active"], dx_t=d["dx_t"],
                            dy_t=d["dy_t"], dt=a.dt, theta=a.theta)
    qe, qw = ale.face_thickness_axis(d["h"], d["H"], d["active"], True, 0)
    qn, qs = ale.face_thickness_axis(d["h"], d["H"], d["active"], True, 1)
    out = lt(d["h"], d["u_new"], d["u_old"], d["v_new"], d["v_old"], d["wet_u"], d["wet_v"], d["wu"], d["wv"],
             d["Du"], d["Dv"], qe, qw, qn, qs, theta_bc=a.theta_bc)
    G1 = ops.advect_fct(_d(C0), out.V0, out.V1, out.Qx, out.Qy, out.ok, a.dt)
    G2 = ops.diffuse_h(G1, G1, out.h_dagger, _d(act.astype(np.uint8)), _d(cx), _d(cy), _d(okx), _d(oky), out.V1,
                       100.0, a.dt)
    dz = torch.clamp(out.h_dagger, min=1e-6).reshape(nz, M)
    ops.vertical_implicit(G2.view(nt, nz, M), dz, _d(nu), None, a.dt, out=G2.view(nt, nz, M))
    vc = vertical.VerticalCoordinate(H=d["H"], active=d["active"], rule="zstar", params={"dz_top": dz_top})
    h_new, G3 = vc.regrid_remap(_d(eta), out.h_dagger, G2)
    _eq(out.h_dagger, hd, "h_dagger")
    _eq(G1, C1, "advected stack")
    _eq(G2, C2, "diffused stack")
    _eq(h_new, h_ref, "h_new")
    _eq(G3, C_ref, "remapped stack")
    assert fl["n_limited"] > 0 or shape[0] < 100
