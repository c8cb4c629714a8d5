"""
GPU tests of the vertical regrid and remap: the HIP kernels are bit-identical to the oracle's C restatement of the
reference arithmetic on columns of every kind (thick, z*-like, interleaved vanished layers, 1-3 active cells, dry,
round-off thickness mismatches, vanished targets), for PCM, PLM and PPM_H4, equal and different layer counts, a
field stack and single fields; the z* and sigma rules on a regional grid; and, after the ALE layer transport, the
regrid-remap of a ten-tracer stack.  At 1024 x 1024 x 50 the column contents are checked to round-off.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import cref
from oracle import numpy_ref as R
from pyocean_amd import ale, synth, vertical
from tests import cases

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu case]", *a, flush=True)


def _d(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _eq(gpu_t, ref, what):
    g = gpu_t.cpu().numpy()
    assert g.shape == ref.shape, (what, g.shape, ref.shape)
    bad = ~((g == ref) | (np.isnan(g) & np.isnan(ref)))
    if bad.any():
        w = tuple(np.argwhere(bad)[0])
        raise AssertionError(f"{what}: {int(bad.sum())} of {g.size} differ, first at {w}: gpu {g[w]!r} ref {ref[w]!r}")


@pytest.mark.parametrize("scheme", [vertical.PCM, vertical.PLM, vertical.PPM_H4])
@pytest.mark.parametrize("n0,n1", [(50, 50), (30, 45), (40, 12), (4, 4), (60, 140)])
def test_remap_matches_the_oracle_bit_for_bit(gpu, scheme, n0, n1):
    _say("remap", scheme, n0, n1)
    hs, f, ht, wet = cases.remap_case(10 + n0 + n1, n0=n0, n1=n1)
    out = vertical.remap_columns(_d(hs), _d(f), _d(ht), scheme, _d(wet))
    _eq(out, cref.remap_columns(hs, f, ht, scheme, wet), "stack")
    one = vertical.remap_columns(_d(hs), _d(f[1]), _d(ht), scheme, _d(wet))
    _eq(one, cref.remap_columns(hs, f[1], ht, scheme, wet), "single field")


@pytest.mark.parametrize("nf,n1", [(2, 20), (10, 20), (14, 20), (17, 20), (20, 20), (10, 100)])
def test_remap_of_many_fields_matches_the_oracle(gpu, nf, n1):
    # the fields go two to a thread and the field groups of a column block to rows of workgroups that no group
    # leaves idle (10 fields: 5 groups, a row of 3 waves and a row of 2, two launches; 17 fields: 9 groups, the last
    # holding one field, three rows of 3; 14 fields: a row of 4 and a row of 3; 20 fields: a row of 4 and two of 3);
    # each workgroup stages its 64 columns' target thicknesses in LDS (with 20 and 100 target layers; the 140-layer
    # targets above take them from HBM)
    _say("remap fields", nf, n1)
    hs, f, ht, wet = cases.remap_case(77 + nf, n0=24, n1=n1, nf=nf)
    out = vertical.remap_columns(_d(hs), _d(f), _d(ht), vertical.PPM_H4, _d(wet))
    _eq(out, cref.remap_columns(hs, f, ht, vertical.PPM_H4, wet), f"{nf} fields")


def test_remap_known_answers(gpu):
    _say("remap known answers")
    h0 = np.array([0, 1, 0, 0, 2, 3, 1], float)[:, None, None]
    u0 = np.array([5, 10, 11, 12, 13, 15, 16], float)[:, None, None]
    out = vertical.remap_columns(_d(h0), _d(u0), _d(np.ones((7, 1, 1))), vertical.PPM_H4, _d(np.ones((1, 1), bool)))
    np.testing.assert_allclose(out.cpu().numpy()[:, 0, 0], [10, 12.642, 13.358, 14.338, 15.052, 15.61, 16], atol=1e-3)
    u = np.array([20.0, 21.0, 22.0, 23.0])[:, None, None]
    for scheme in (vertical.PCM, vertical.PLM, vertical.PPM_H4):
        z = vertical.remap_columns(_d(np.zeros((4, 1, 1))), _d(u), _d(np.zeros((2, 1, 1))), scheme,
                                   _d(np.ones((1, 1), bool)))
        np.testing.assert_allclose(z.cpu().numpy()[:, 0, 0], 21.5)


@pytest.mark.parametrize("rule", ["zstar", "sigma"])
def test_regrid_matches_the_oracle_bit_for_bit(gpu, rule):
    _say("regrid", rule)
    a = synth.ale_inputs(96, 70, 50, seed=3)
    rng = np.random.default_rng(4)
    eta = rng.uniform(-0.8, 0.8, a.H.shape) * a.active
    h = a.h * rng.uniform(0.9, 1.1, a.h.shape)
    dz = np.full(50, 40.0) if rule == "zstar" else np.full(50, 1.0 / 50)
    params = {"dz_top": dz} if rule == "zstar" else {"dsigma": dz}
    H = a.H.copy()
    if rule == "sigma":
        H[~a.active] = np.nan                                          # NaN depth reads as 0
    h_new, wet = vertical.regrid(_d(H), _d(eta), _d(a.active), _d(h), rule, params)
    r_new, r_wet = cref.regrid(H, eta, a.active, h, rule, dz)
    _eq(wet, r_wet.astype(np.uint8), "wet")
    _eq(h_new, r_new, "h_new")


def test_regrid_rejects_a_wet_column_it_cannot_build(gpu):
    _say("regrid errors")
    H = np.full((3, 4), 100.0)
    H[1, 2] = -5.0
    eta = np.full((3, 4), 10.0)                                         # D > 0 everywhere: column (1, 2) is wet
    h = np.full((5, 3, 4), 2.0)
    with pytest.raises(ValueError, match="j=1, i=2"):
        vertical.regrid(_d(H), _d(eta), _d(np.ones((3, 4), bool)), _d(h), "zstar", {"dz_top": np.full(5, 20.0)})
    with pytest.raises(ValueError, match="rho_target"):
        vertical.regrid(_d(H), _d(eta), _d(np.ones((3, 4), bool)), _d(h), "hybrid", {"dz_top": np.ones(5)})
    rho = np.full((5, 3, 4), 1025.0)
    params = {"dz_top": np.full(5, 20.0), "rho_target": np.linspace(1024.0, 1026.0, 6)}
    with pytest.raises(ValueError, match="j=1, i=2"):                 # H <= 0 without a sigma zone
        vertical.regrid(_d(H), _d(eta), _d(np.ones((3, 4), bool)), _d(h), "hybrid", params, rho=_d(rho))
    H[1, 2] = 100.0
    rho[3, 2, 1] = np.nan
    with pytest.raises(ValueError, match="j=2, i=1"):                 # a non-finite density in a wet column
        vertical.regrid(_d(H), _d(eta), _d(np.ones((3, 4), bool)), _d(h), "hybrid", params, rho=_d(rho))


def test_regrid_remap_after_the_layer_transport_matches_the_oracle(gpu):
    _say("layer transport -> regrid_remap")
    a = synth.ale_inputs(64, 48, 20, seed=5)
    r = R.ale_layer_transport(a.h, a.H, a.active, a.area, a.dx_t, a.dy_t, a.Lu, a.Lv, a.u_new, a.u_old, a.v_new,
                              a.v_old, a.wet_u, a.wet_v, a.wu, a.wv, a.Du, a.Dv, a.dt, a.theta, a.theta_bc)
    h_dag = r["h_dagger"]
    rng = np.random.default_rng(6)
    eta = h_dag.sum(0) - a.H
    C = rng.normal(10.0, 2.0, (10,) + h_dag.shape)
    dz = np.full(20, 2000.0 / 20)
    vc = vertical.VerticalCoordinate(H=_d(a.H), active=_d(a.active), rule="zstar", params={"dz_top": dz})
    h_new, Cn = vc.regrid_remap(_d(eta), _d(h_dag), _d(C))
    r_new, r_wet = cref.regrid(a.H, eta, a.active, h_dag, "zstar", dz)
    _eq(h_new, r_new, "h_new")
    _eq(Cn, cref.remap_columns(h_dag, C, r_new, cref.PPM_H4, r_wet), "C remapped")
    h_new2, lst = vc.regrid_remap(_d(eta), _d(h_dag), [_d(C[0]), _d(C[3])])
    _eq(lst[1], cref.remap_columns(h_dag, C[3], r_new, cref.PPM_H4, r_wet), "list field")


def test_remap_at_full_size_conserves_every_column(gpu):
    _say("remap 1024x1024x50")
    a = synth.ale_inputs(1024, 1024, 50, seed=0)
    rng = np.random.default_rng(1)
    hs = _d(a.h)
    ht = _d(a.h * rng.uniform(0.8, 1.2, a.h.shape))
    ht = ht * (hs.sum(0) / torch.where(ht.sum(0) > 0, ht.sum(0), torch.ones_like(ht.sum(0))))[None]
    wet = _d(a.active & (a.h.sum(0) > 0))
    C = torch.empty((4,) + a.h.shape, dtype=torch.float64, device="cuda")
    C[0] = 12.5
    C[1:] = _d(rng.normal(10.0, 2.0, (3,) + a.h.shape))
    out = vertical.remap_columns(hs, C, ht, vertical.PPM_H4, wet)
    w = wet.bool()
    # a uniform field stays uniform to round-off: a piece's length is a difference of interface heights of up to the
    # column depth D, so a thin target cell carries a relative error of order eps D / h
    D = ht.sum(0)[None]
    tol = 64 * 2.3e-16 * 12.5 * (D / torch.where(ht > 0, ht, torch.ones_like(ht)))
    assert ((out[0] - 12.5).abs() <= tol)[:, w].all()
    c0 = (hs[None] * C).sum(1)[:, w]
    c1 = (ht[None] * out).sum(1)[:, w]
    assert ((c1 - c0).abs() <= 1e-12 * (hs[None] * C.abs()).sum(1)[:, w] + 1e-300).all()
    assert torch.isfinite(out[:, :, w]).all()


def _hybrid_case(seed, n):
    rng = np.random.default_rng(seed)
    ny, nx = 13, 17
    H = rng.uniform(3.0, 400.0, (ny, nx))
    H[0, :3] = [-0.5, 0.0, 1.0]
    eta = rng.uniform(-0.3, 0.3, (ny, nx))
    eta[0, 0] = 0.8
    act = rng.random((ny, nx)) > 0.1
    frac = rng.uniform(0.0, 1.0, (n, ny, nx))
    frac[rng.random(frac.shape) < 0.3] = 0.0
    tot = frac.sum(0)
    tot[tot == 0] = 1.0
    h = np.where((H + eta)[None] > 0, frac / tot * (H + eta)[None], 0.0)
    h = np.where(h < 2e-3, rng.choice([0.0, 1e-3, 1.5e-3], h.shape), h)
    zc = -np.cumsum(h[::-1], axis=0)[::-1] + 0.5 * h
    rho = 1024.0 - 0.01 * zc + rng.normal(0, 0.2, h.shape)
    targets = np.sort(rng.uniform(1023.0, 1030.0, n + 1))
    return H, eta, act, h, rho, targets, rng


@pytest.mark.parametrize("keep", [True, False])
@pytest.mark.parametrize("window", [False, True])
@pytest.mark.parametrize("n", [10, 50])
def test_hybrid_regrid_matches_the_oracle_bit_for_bit(gpu, keep, window, n):
    _say("regrid hybrid", keep, window, n)
    H, eta, act, h, rho, targets, rng = _hybrid_case(5 + n + 2 * keep + window, n)
    params = dict(dz_top=rng.uniform(1.0, 10.0, n), dsigma=np.full(n, 1.0 / n), rho_target=targets,
                  max_h_top=np.where(rng.random(n) < 0.3, 30.0, np.inf), d_sigma=20.0,
                  d_hybrid=100.0 if window else 0.0, keep_isopycnal=keep)
    if not window:
        H = np.where(H > 0, H, 5.0)
    h_new, wet = vertical.regrid(_d(H), _d(eta), _d(act), _d(h), "hybrid", params, rho=_d(rho))
    r_new, r_wet = cref.regrid_hybrid(H, eta, act, h, rho, params)
    _eq(wet, r_wet.astype(np.uint8), "wet")
    _eq(h_new, r_new, "h_new")


def test_hybrid_regrid_remap_of_a_stack_matches_the_oracle(gpu):
    _say("regrid_remap hybrid")
    n = 30
    H, eta, act, h, rho, targets, rng = _hybrid_case(9, n)
    H = np.where(H > 0, H, 5.0)
    params = dict(dz_top=np.full(n, 2.0), rho_target=targets)
    C = rng.normal(10.0, 1.0, (4, n) + H.shape)
    vc = vertical.VerticalCoordinate(H=_d(H), active=_d(act), rule="hybrid", params=params)
    h_new, Cn = vc.regrid_remap(_d(eta), _d(h), _d(C), rho=_d(rho))
    r_new, r_wet = cref.regrid_hybrid(H, eta, act, h, rho, params)
    _eq(h_new, r_new, "h_new")
    _eq(Cn, cref.remap_columns(h, C, r_new, cref.PPM_H4, r_wet), "C remapped")


_EOS = np.load(__import__("pathlib").Path(__file__).parent / "golden" / "mom6_eos_wright.npz")


def test_densities_match_the_oracle_bit_for_bit(gpu):
    _say("regrid densities")
    from pyocean_amd import eos
    T, S, p = _EOS["T"], _EOS["S"], _EOS["p"]
    w = eos.WrightFullEOS()
    _eq(w.density(_d(T), _d(S), _d(p)), R.wright_density(T, S, p), "wright rho(T, S, p)")
    _eq(w.potential_density(_d(T), _d(S), 2e7), R.wright_density(T, S, 2e7), "wright sigma_2 + 1000")
    _eq(w.density(_d(T), _d(S)), R.wright_density(T, S, 0.0), "wright surface")
    np.testing.assert_allclose(w.density(_d(T), _d(S), _d(p)).cpu().numpy(), _EOS["w_rho"], rtol=1e-12, atol=0)
    lin = eos.make_eos("linear", rho0=1025.0, t_coef=0.21, t_ref=10.0, s_coef=0.78, s_ref=35.0)
    _eq(lin.density(_d(T), _d(S)), R.linear_density(T, S, 1025.0, 0.21, 10.0, 0.78, 35.0), "linear rho(T, S)")
    _eq(lin.density(_d(T)), R.linear_density(T, None, 1025.0, 0.21, 10.0, 0.78, 35.0), "linear rho(T)")
    T3 = _d(T[:2040].reshape(8, 15, 17))
    assert lin.density(T3).shape == (8, 15, 17)


def test_hybrid_regrid_remap_follows_the_equation_of_state(gpu):
    _say("regrid_remap hybrid from T and S")
    from pyocean_amd import eos
    n = 20
    H, eta, act, h, _, _, rng = _hybrid_case(13, n)
    H = np.where(H > 0, H, 5.0)
    zc = -np.cumsum(h[::-1], axis=0)[::-1] + 0.5 * h
    T = 4.0 + 16.0 * np.exp(zc / 150.0) + rng.normal(0, 0.1, h.shape)
    S = 35.0 + 0.002 * zc
    params = dict(dz_top=np.full(n, 2.0), rho_target=np.linspace(1031.0, 1037.5, n + 1), dsigma=np.full(n, 1.0 / n),
                  d_sigma=10.0, d_hybrid=40.0)
    C = rng.normal(10.0, 1.0, (3, n) + H.shape)
    for e, rho in ((eos.WrightFullEOS(), R.wright_density(T, S, 2e7)),
                   (eos.LinearEOS(t_coef=0.2, s_coef=0.76), R.linear_density(T, S, 1000.0, 0.2, 0.0, 0.76, 35.0))):
        tg = params["rho_target"] if not e.linear else np.linspace(rho.min(), rho.max(), n + 1)
        pr = dict(params, rho_target=tg)
        vc = vertical.VerticalCoordinate(H=_d(H), active=_d(act), rule="hybrid", params=pr, eos=e)
        _eq(vc.regrid_density(_d(T), _d(S)), rho, f"{e.name} regrid_density")
        h_new, Cn = vc.regrid_remap(_d(eta), _d(h), _d(C), T=_d(T), S=_d(S))
        r_new, r_wet = cref.regrid_hybrid(H, eta, act, h, rho, pr)
        _eq(h_new, r_new, f"{e.name} h_new")
        _eq(vc.regrid(_d(eta), _d(h), T=_d(T), S=_d(S)), r_new, f"{e.name} regrid")
        _eq(Cn, cref.remap_columns(h, C, r_new, cref.PPM_H4, r_wet), f"{e.name} C remapped")


@pytest.mark.parametrize("seed", range(8))
def test_hybrid_positions_match_the_oracle_on_exact_hits_and_odd_targets(gpu, seed):
    # the kernel finds the target positions by bisection on monotone edges, where the reference scans: densities in
    # constant blocks make the edges equal to the block values, which the targets then hit exactly; some targets are
    # unsorted, NaN or outside the range, some sources noisy (non-monotone before the running minimum) or unstratified
    _say("regrid hybrid positions", seed)
    rng = np.random.default_rng(100 + seed)
    n = int(rng.choice([2, 3, 7, 20, 50]))
    ny, nx = 11, 13
    H = rng.uniform(5.0, 300.0, (ny, nx))
    eta = rng.uniform(-0.2, 0.2, (ny, nx))
    act = rng.random((ny, nx)) > 0.1
    frac = rng.uniform(0, 1, (n, ny, nx))
    frac[rng.random(frac.shape) < 0.3] = 0.0
    tot = frac.sum(0)
    tot[tot == 0] = 1.0
    h = frac / tot * (H + eta)[None]
    h = np.where(h < 2e-3, rng.choice([0.0, 1e-3, 1.5e-3], h.shape), h)
    vals = np.sort(rng.choice([1024.0, 1025.0, 1025.5, 1026.0, 1027.25], 4))[::-1]
    rho = vals[np.sort(rng.integers(0, 4, n))][:, None, None] * np.ones((1, ny, nx))
    if seed % 3 == 0:
        rho = rho + rng.normal(0, 0.3, rho.shape)
    rho[:, 0, 0] = 1025.0
    cand = np.concatenate([vals, 0.5 * (vals[1:] + vals[:-1]), rng.uniform(1023, 1028, 8), [1020.0, 1030.0]])
    targets = np.sort(rng.choice(cand, n + 1))
    if seed % 4 == 1:
        targets = rng.permutation(targets)
    if seed % 4 == 2:
        targets[n // 2] = np.nan
    params = dict(dz_top=rng.uniform(0.5, 5.0, n), dsigma=np.full(n, 1.0 / n), rho_target=targets,
                  max_h_top=np.where(rng.random(n) < 0.3, 30.0, np.inf), d_sigma=10.0,
                  d_hybrid=60.0 if seed % 2 else 0.0, keep_isopycnal=bool(seed % 3))
    h_new, wet = vertical.regrid(_d(H), _d(eta), _d(act), _d(h), "hybrid", params, rho=_d(rho))
    r_new, r_wet = cref.regrid_hybrid(H, eta, act, h, rho, params)
    _eq(wet, r_wet.astype(np.uint8), "wet")
    _eq(h_new, r_new, "h_new")
