"""
GPU tests of the serving loop: jobs that join and leave a running batch (by free slots, by the HBM budget and by
arrival time) each end bit-identical to the same job run alone, and their content sums are the response.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from pyocean_amd import serve, synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu case]", *a, flush=True)


SHAPES = [(80, 96, 5, 2), (130, 70, 4, 3), (100, 100, 6, 2)]


@pytest.fixture(scope="module")
def cases_():
    return [synth.regional(nx=nx, ny=ny, nz=nz, nt=nt, seed=60 + m) for m, (nx, ny, nz, nt) in enumerate(SHAPES)]


@pytest.fixture(scope="module")
def grids(cases_):
    return [serve.Grid.from_case(c) for c in cases_]


def _oracle(case, steps, memo={}):                                       # noqa: B006 (a per-process cache)
    key = (case.info["name"], float(case.dt), steps)
    if key not in memo:
        from oracle import cref
        c = case.C
        for _ in range(steps):
            c = cref.advect_fct(c, case.V0, case.V1, case.Qx, case.Qy, case.ok, case.dt)
        memo[key] = c
    return memo[key]


def _check(grids, jobs, precision, cases_):
    """Each job equals its solo run bit for bit and returns its content; against the oracle stepped as often, an
    FP64 job is bit-identical and a split job stays bounded and within 1e-5 of each tracer's span per step."""
    for j in jobs:
        alone = serve.advance_alone(grids[j.grid], j.steps, precision)
        assert torch.equal(j.state, alone), (j.id, j.grid, j.steps)
        V1 = grids[j.grid].V1
        ref = (alone * V1).sum(dim=(1, 2, 3)).cpu().numpy()
        np.testing.assert_allclose(j.content, ref, rtol=1e-12)
        case = cases_[j.grid]
        want = _oracle(case, j.steps)
        got = j.state.cpu().numpy()
        if precision == "f64":
            assert np.array_equal(got, want), (j.id, j.grid, j.steps)
            continue
        okb = np.broadcast_to(case.ok, case.C.shape)
        for t in range(case.C.shape[0]):
            lo, hi = case.C[t][okb[t]].min(), case.C[t][okb[t]].max()
            span = hi - lo
            assert np.abs(got[t] - want[t])[okb[t]].max() <= 1e-5 * span * j.steps, (j.id, t)
            assert got[t][okb[t]].min() >= lo - 1e-12 * span and got[t][okb[t]].max() <= hi + 1e-12 * span


@pytest.mark.parametrize("precision", ["split", "f64"])
def test_jobs_joining_by_free_slots_equal_their_solo_runs(gpu, grids, cases_, precision):
    _say("serve slots", precision)
    steps = [6, 2, 3, 5, 1, 4]
    jobs = [serve.Job(id=i, grid=i % 3, steps=s, arrival=0.0) for i, s in enumerate(steps)]
    stats = serve.Server(grids, hbm_budget=64e9, max_active=2, precision=precision).run(jobs, keep_states=True)
    _check(grids, jobs, precision, cases_)
    # continuous batching: some job joined a batch whose other member had started earlier and was still running
    joined = [(a, b) for a in jobs for b in jobs
              if b.admit_step > a.admit_step and b.admit_step <= a.done_step < b.done_step]
    assert joined
    assert stats["batched_steps"] < sum(steps)
    assert max(j.done_step - j.admit_step + 1 for j in jobs) == max(steps)
    for j in jobs:
        assert j.done_step - j.admit_step + 1 == j.steps


def test_admission_waits_for_memory_and_keeps_arrival_order(gpu, grids, cases_):
    _say("serve budget")
    need = [g.job_bytes("split") for g in grids]
    resident = sum(g.resident_bytes() for g in grids)
    budget = resident + need[0] + need[1] + 1                # two of the first grids fit, never three
    jobs = [serve.Job(id=i, grid=i % 2, steps=2 + i % 3, arrival=0.0) for i in range(5)]
    srv = serve.Server(grids, hbm_budget=budget, max_active=8, precision="split")
    srv.run(jobs, keep_states=True)
    _check(grids, jobs, "split", cases_)
    admitted = sorted(jobs, key=lambda j: (j.admit_step, j.id))
    assert [j.id for j in admitted] == [j.id for j in jobs]  # first come, first served
    for s in range(max(j.done_step for j in jobs) + 1):
        live = [j for j in jobs if j.admit_step <= s <= j.done_step]
        assert sum(need[j.grid] for j in live) <= budget - resident
    with pytest.raises(RuntimeError, match="more than the whole budget"):
        serve.Server(grids, hbm_budget=resident + need[0] // 2, precision="split").run(
            [serve.Job(id=0, grid=0, steps=1, arrival=0.0)])


def test_poisson_arrivals_are_served_and_timed(gpu, grids, cases_):
    _say("serve poisson")
    jobs = serve.poisson_jobs(12, rate=400.0, n_grids=3, steps_lo=1, steps_hi=8, big_grid=2, big_prob=0.2, seed=3)
    stats = serve.Server(grids, hbm_budget=64e9, max_active=4, precision="split").run(jobs, keep_states=True)
    _check(grids, jobs, "split", cases_)
    for j in jobs:
        assert j.done >= j.admitted >= j.arrival
    assert stats["p99_latency_s"] >= stats["p50_latency_s"] > 0
    assert stats["updates"] == sum(j.steps * grids[j.grid].C0.numel() for j in jobs)
