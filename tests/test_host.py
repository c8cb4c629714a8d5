"""Host-side logic of the drop-in component and the workload generators (CPU only, no kernel launches)."""
from __future__ import annotations

import numpy as np
import pytest

from pyocean_amd import Periodic, synth
from pyocean_amd.contract import Role
from pyocean_amd.tracers import OpenEdges, TracerAdvection


def _adv(**kw):
    return TracerAdvection(scheme=kw.get("scheme", "weno5"), time=kw.get("time", "ssprk3"), dt=30.0,
                           ob=OpenEdges(kw.get("open_", ())))


def test_component_contract_matches_the_reference():
    adv = _adv()
    assert adv.role == Role.TRACER_ADVECTION and adv.name == "tracer_advection"
    assert adv.reads == ("T", "tr", "Qx", "Qy", "Fz") and adv.writes == ("T", "tr")
    assert adv.check(None) == [] and adv.state_fields() == {}
    assert adv.record_open_flux is False and adv.last_open_flux is None
    assert not adv.per


def test_volume_shape_mismatch_raises_value_error_before_the_gpu():
    adv = _adv()
    C = np.zeros((2, 3, 4, 5))
    with pytest.raises(ValueError, match="volume shape"):
        adv.advect(C, np.zeros((3, 4, 6)), np.zeros((3, 4, 5)), np.zeros((3, 4, 6)), np.zeros((3, 5, 5)),
                   np.zeros((4, 4, 5)), np.ones((3, 4, 5), bool))
    with pytest.raises(ValueError, match="Qx"):
        adv.advect_fct(C, np.zeros((3, 4, 5)), np.zeros((3, 4, 5)), np.zeros((3, 4, 5)), np.zeros((3, 5, 5)),
                       np.ones((3, 4, 5), bool))


def test_empty_stack_gives_empty_results():
    """A stack with no tracers: the reference's loops run over no planes and return empty arrays of the input's
    shape (the face column (0, N, ny)); the open-edge budget is empty too.  Nothing reaches the device."""
    adv = _adv()
    adv.record_open_flux = True
    C = np.zeros((0, 3, 4, 5))
    V, Qx, Qy, Fz, ok = (np.ones((3, 4, 5)), np.zeros((3, 4, 6)), np.zeros((3, 5, 5)), np.zeros((4, 4, 5)),
                         np.ones((3, 4, 5), bool))
    out, fx = adv.advect_fct(C, V, V, Qx, Qy, ok, flux_face_i=2)
    assert out.shape == C.shape and fx.shape == (0, 3, 4)
    assert adv.last_open_flux.shape == (0,)
    assert adv.advect(C, V, V, Qx, Qy, Fz, ok).shape == C.shape
    out, fx = adv.advect(C, V, V, Qx, Qy, Fz, ok, flux_face_i=0)
    assert out.shape == C.shape and fx.shape == (0, 3, 4)


@pytest.mark.parametrize("scheme,rk_bits,fct_bits", [("weno5", 0, 0), ("weno5js", 16, 16), ("upwind3", 64, 16),
                                                     ("upwind1", 128, 16)])
def test_face_flags_follow_the_reference_dispatch(scheme, rk_bits, fct_bits):
    """The RK3 step takes UP3 / donor faces for the upwind schemes (tracers.py:110-115); the FCT step and faces_h
    run WENO5-JS for every scheme but 'weno5' (tracers.py:166-176)."""
    adv = _adv(scheme=scheme, open_=("west", "north"))
    assert adv._flags(rk=True) == rk_bits | 1 | 8
    assert adv._flags() == fct_bits | 1 | 8
    assert adv._flags(unlimited=True) == fct_bits | 1 | 8 | 32


def test_flag_word_rejects_unknown_up_orders_and_a_missing_fz():
    from pyocean_amd import ops
    assert ops.flags_of(2, up=3) == 66 and ops.flags_of(0, up=1) == 128
    with pytest.raises(ValueError, match="up = 2"):
        ops.flags_of(0, up=2)
    with pytest.raises(ValueError, match="need Fz"):
        ops._up_faces(64, None)
    assert ops._up_faces(16, None) is False


def test_numpy_twin_names_are_the_gpu_methods():
    """advect_fct_numpy / advect_numpy / rhs_numpy keep the reference's names and signatures (tracers.py:139, 256,
    303) and call the GPU methods: no CPU path behind them."""
    import inspect
    adv = _adv()
    for twin, gpu in (("advect_fct_numpy", "advect_fct"), ("advect_numpy", "advect"), ("rhs_numpy", "rhs")):
        assert list(inspect.signature(getattr(adv, twin)).parameters) == \
            list(inspect.signature(getattr(adv, gpu)).parameters)
    C = np.zeros((1, 2, 3, 4))
    with pytest.raises(ValueError, match="volume shape"):
        adv.advect_numpy(C, np.zeros((2, 3, 5)), np.zeros((2, 3, 4)), np.zeros((2, 3, 5)), np.zeros((2, 4, 4)),
                         np.zeros((3, 3, 4)), np.ones((2, 3, 4), bool))


def test_centred_faces_flag_word():
    """rhs / flux_x with centred=True take the centred faces whatever the scheme (tracers.py:110-113)."""
    for scheme in ("weno5", "upwind3"):
        assert _adv(scheme=scheme, open_=("east",))._flags(centred=True) == 256 | 2
    from pyocean_amd import ops
    assert ops.flags_of(5, centred=True) == 261
    with pytest.raises(ValueError, match="need Fz"):
        ops._up_faces(256, None)


def test_periodic_axes():
    assert not Periodic() and Periodic("x").x and not Periodic("x").y
    assert Periodic(("y", "x")).axes == ("x", "y")
    with pytest.raises(ValueError, match="unknown periodic axes"):
        Periodic(("z",))


def test_solid_body_case_is_non_divergent_and_closed():
    c = synth.solid_body(n=64, nz=4)
    div = c.Qx[..., 1:] - c.Qx[..., :-1] + c.Qy[:, 1:] - c.Qy[:, :-1]
    assert np.abs(div).max() <= 1e-9 * np.abs(c.Qx).max()
    assert not c.Qx[..., [0, -1]].any() and not c.Qy[:, [0, -1]].any()
    np.testing.assert_allclose(c.V1, c.V0, rtol=1e-13)
    ux = np.abs(c.Qx).max() / (200.0 / 4 * 1000.0)
    assert abs(ux * c.dt / 1000.0 - 0.4) < 1e-12                        # Courant 0.4
    assert set(np.unique(c.C[1])) == {0.0, 1.0}


def test_regional_case_is_valid_for_the_limiter():
    c = synth.regional(nx=96, ny=80, nz=10, nt=4, seed=5)
    assert 0.1 < c.info["land"] < 0.3
    assert c.info["max_outflow_fraction"] < 1.0                          # limit_outflow's condition
    assert (c.V1[c.ok] > 0).all() and not c.ok[c.V0 <= 0].any()
    land = c.V0[0] <= 0
    assert not c.Qx[:, :, 1:][:, land].any() and not c.Qx[:, :, :-1][:, land].any()
    assert not c.Qx[..., [0, -1]].any() and not c.Qy[:, [0, -1]].any()
    s = synth.regional(nx=96, ny=80, nz=10, nt=1, seed=5, sigma=True)
    div = s.Qx[..., 1:] - s.Qx[..., :-1] + s.Qy[:, 1:] - s.Qy[:, :-1] + s.Fz[1:] - s.Fz[:-1]
    np.testing.assert_allclose(s.V1, s.V0 - s.dt * div, rtol=0, atol=1e-6 * s.V0.max())
    assert not s.Fz[0].any() and not s.Fz[-1].any() and np.abs(s.Fz).max() > 0


def test_batch_shapes_follow_the_batched_case():
    shapes = synth.batch_shapes(64)
    assert len(shapes) == 64
    for nx, ny, nz in shapes:
        assert 96 <= nx <= 768 and 96 <= ny <= 768 and nz in (20, 30, 50, 75)


def test_poisson_job_stream_follows_the_serving_case():
    from pyocean_amd import serve
    jobs = serve.poisson_jobs(4000, rate=20.0, n_grids=65, big_grid=64, big_prob=0.02, seed=7)
    t = np.array([j.arrival for j in jobs])
    steps = np.array([j.steps for j in jobs])
    grids = np.array([j.grid for j in jobs])
    assert np.all(np.diff(t) >= 0) and abs(len(jobs) / t[-1] - 20.0) < 1.0          # mean rate
    gaps = np.diff(t)
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.1                                 # exponential gaps: CV = 1
    assert steps.min() >= 10 and steps.max() <= 500
    assert abs(np.median(np.log(steps)) - 0.5 * (np.log(10) + np.log(500))) < 0.15   # log-uniform
    assert 0.01 < np.mean(grids == 64) < 0.03 and set(grids[grids != 64]) == set(range(64))
    again = serve.poisson_jobs(4000, rate=20.0, n_grids=65, big_grid=64, big_prob=0.02, seed=7)
    assert [(j.arrival, j.steps, j.grid) for j in again] == [(j.arrival, j.steps, j.grid) for j in jobs]


def test_next_row_mirrors_reject_bad_arguments_before_any_kernel():
    # the ALE, freshwater and vertical-coordinate mirrors check names, shapes and devices on the host and raise
    # the reference's ValueError, with no GPU call (CPU tensors are refused as such)
    import torch

    from pyocean_amd import ale, freshwater, vertical
    cpu = torch.zeros((2, 3, 4), dtype=torch.float64)
    with pytest.raises(ValueError, match="thickness_flux"):
        ale.LayerTransport(Lu=None, Lv=None, area=None, active=None, dx_t=None, dy_t=None, dt=1.0, theta=0.5,
                           thickness_flux="weno")
    with pytest.raises(ValueError, match="unknown open edge"):
        ale.open_face_transports(None, None, cpu, None, None, None, None, ("up",), None, None, None, None, None,
                                 None, 0.5)
    assert ale.open_face_transports("Qx", "Qy", cpu, None, None, None, None, (), None, None, None, None, None,
                                    None, 0.5) == ("Qx", "Qy")
    with pytest.raises(ValueError, match="GPU tensor"):
        freshwater.freshwater_volumes(cpu, torch.zeros((3, 4), dtype=torch.float64), 1e-6)
    with pytest.raises(ValueError, match="contiguous float64 GPU tensor"):
        freshwater.apply_freshwater(torch.zeros((2, 2, 3, 4), dtype=torch.float64), {"T": 0, "n": 2}, cpu, cpu,
                                    torch.zeros((3, 4), dtype=torch.float64), dt=1.0)
    with pytest.raises(ValueError, match="unknown rule"):
        vertical.VerticalCoordinate(H=None, active=None, rule="isopycnal", params={})
    with pytest.raises(ValueError, match="needs params"):
        vertical.regrid(None, None, None, cpu, "hybrid", {"dz_top": [1.0]})
    with pytest.raises(ValueError, match="layer densities rho"):
        vertical.regrid(None, None, None, cpu, "hybrid", {"dz_top": [1.0], "rho_target": [1.0]})


def test_stack_layout_matches_the_reference_cases():
    """The positions of [T, (T²), S, (S²), tr…] for the model switches of the reference's own layout tests
    (test_n2a_salinity.py:188-195, test_n2a_tef.py:247-249; the reference model's mix_diagnostic defaults to on)."""
    from types import SimpleNamespace as M

    from pyocean_amd.stack import stack_layout
    assert stack_layout(M(ale=False)) == {"T": 0, "T2": 1, "n": 2}
    assert stack_layout(M(ale=False, salinity=True)) == {"T": 0, "T2": 1, "S": 2, "n": 3}
    assert stack_layout(M(ale=True, mix_diagnostic=True, n_tracers=2)) == {"T": 0, "T2": 1, "tr": 2, "n": 4}
    assert stack_layout(M(ale=True, salinity=True, mix_diagnostic=False, n_tracers=1)) == \
        {"T": 0, "S": 1, "tr": 2, "n": 3}
    assert stack_layout(M(ale=True, salinity=True, mix_diagnostic=True, n_tracers=1)) == \
        {"T": 0, "T2": 1, "S": 2, "tr": 3, "n": 4}
    assert "S2" in stack_layout(M(ale=False, salinity=True, variance_budget=False, tef_face=(3, "x")))
    # the passive tracers count only in ALE; the variance budget keeps T² and adds S²
    assert stack_layout(M(ale=False, n_tracers=3)) == {"T": 0, "T2": 1, "n": 2}
    assert stack_layout(M(ale=True, salinity=True, variance_budget=True)) == {"T": 0, "T2": 1, "S": 2, "S2": 3, "n": 4}


def test_inflow_stack_follows_the_layout_on_host_and_torch_arrays():
    """OpenBoundary.inflow_stack (open_boundary.py:353-380): the stack itself without open edges or inflow; with a
    layout the inflow fields in its order (squares as x·x); without one [T, (T²), tr…] by the stack's length."""
    from types import SimpleNamespace as M

    import torch

    from pyocean_amd.stack import inflow_stack, stack_layout
    rng = np.random.default_rng(3)
    T, S = rng.normal(10, 2, (4, 5, 6)), rng.normal(35, 1, (4, 5, 6))
    tr = [rng.normal(0, 1, (4, 5, 6)) for _ in range(2)]
    lay = stack_layout(M(ale=True, salinity=True, variance_budget=True, n_tracers=2))
    c = rng.normal(0, 1, (lay["n"], 4, 5, 6))
    shut = M(open=(), ale=True, n_tracers=2)
    assert inflow_stack(shut, c, T, tr, S, lay) is c
    assert inflow_stack(M(open=("west",)), c, None, tr, S, lay) is c
    ob = M(open=("west", "south"), ale=True, n_tracers=2)
    got = inflow_stack(ob, c, T, tr, S, lay)
    want = np.stack([T, T * T, S, S * S, tr[0], tr[1]])
    assert got.flags["C_CONTIGUOUS"] and np.array_equal(got, want)
    tt = inflow_stack(ob, torch.from_numpy(c), torch.from_numpy(T), [torch.from_numpy(a) for a in tr],
                      torch.from_numpy(S), lay)
    assert isinstance(tt, torch.Tensor) and np.array_equal(tt.numpy(), want)
    # no layout: [T, T², tr…] when the stack has room for T², [T, tr…] when it does not
    assert np.array_equal(inflow_stack(ob, np.zeros((4, 4, 5, 6)), T, tr), np.stack([T, T * T, tr[0], tr[1]]))
    assert np.array_equal(inflow_stack(ob, np.zeros((3, 4, 5, 6)), T, tr), np.stack([T, tr[0], tr[1]]))
    assert np.array_equal(inflow_stack(M(open=("north",), ale=False), np.zeros((2, 4, 5, 6)), T, None),
                          np.stack([T, T * T]))
