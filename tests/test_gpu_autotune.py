"""
GPU tests of the autotuner: every launch shape it tries gives bit-identical results (single domain and batch),
the chosen shape is cached and reused, and a tuned batch ends exactly where the same number of untuned steps ends.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from pyocean_amd import autotune, batch, ops, synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu case]", *a, flush=True)


def _dev(case):
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()        # noqa: E731
    return d(case.C), d(case.V0), d(case.V1), d(case.Qx), d(case.Qy), d(case.ok.astype(np.uint8))


@pytest.mark.parametrize("precision", ["f64", "split"])
def test_every_step_launch_shape_gives_the_same_bits(gpu, precision):
    _say("launch shapes", precision)
    case = synth.regional(nx=140, ny=150, nz=4, nt=5, seed=80)
    C, V0, V1, Qx, Qy, ok = _dev(case)
    outs = []
    try:
        for r in autotune.ROWS:
            for w in autotune.WAVES:
                batch.set_tuning(r, w)
                outs.append(((r, w), ops.advect_fct(C, V0, V1, Qx, Qy, ok, case.dt, precision=precision)))
    finally:
        batch.set_tuning(0, 0)
    for shape, o in outs[1:]:
        assert torch.equal(o, outs[0][1]), shape


def test_tune_step_picks_a_shape_and_caches_it(gpu, tmp_path, monkeypatch):
    _say("tune step")
    monkeypatch.setenv("OCX_TUNING_CACHE", str(tmp_path / "t.json"))
    case = synth.regional(nx=130, ny=120, nz=3, nt=3, seed=81)
    C, V0, V1, Qx, Qy, ok = _dev(case)
    C0 = C.clone()
    try:
        best, table = autotune.tune_step(C, V0, V1, Qx, Qy, ok, case.dt, rows=(16, 64), waves=(1, 2), reps=2)
        assert len(table) == 4 and best in [s for s, _ in table]
        assert torch.equal(C, C0)
        again, table2 = autotune.tune_step(C, V0, V1, Qx, Qy, ok, case.dt, rows=(16, 64), waves=(1, 2))
        assert again == best and table2 == []
    finally:
        batch.set_tuning(0, 0)


def test_tuned_batch_matches_untuned_steps(gpu, tmp_path, monkeypatch):
    _say("tune batch")
    monkeypatch.setenv("OCX_TUNING_CACHE", str(tmp_path / "t.json"))
    shapes = [(120, 90, 4, 3, "f64"), (70, 140, 5, 2, "split"), (100, 100, 3, 4, "split")]
    doms, refs = [], []
    for m, (nx, ny, nz, nt, pr) in enumerate(shapes):
        case = synth.regional(nx=nx, ny=ny, nz=nz, nt=nt, seed=90 + m)
        C, V0, V1, Qx, Qy, ok = _dev(case)
        doms.append(batch.Domain(C=C, V0=V0, V1=V1, Qx=Qx, Qy=Qy, ok=ok, dt=case.dt, precision=pr))
        refs.append(C.clone())
    best, table, steps = autotune.tune_batch(doms, rows=(16, 64), waves=(1, 4), groups=(1, 4), reps=1)
    assert len(table) == 8 and steps == 8 * 2 * 2 and steps % 2 == 0
    for m, d in enumerate(doms):
        C = refs[m]
        for _ in range(steps):
            C = ops.advect_fct(C, d.V0, d.V1, d.Qx, d.Qy, d.ok, d.dt, precision=d.precision)
        assert torch.equal(d.C, C), m
    b = batch.Batch(doms, rows_per_chunk=best.rows_per_chunk, waves_per_group=best.waves_per_group,
                    groups_per_cu=best.groups_per_cu)
    b.step()
    b.close()
    assert autotune.tune_batch(doms, rows=(16, 64), waves=(1, 4), groups=(1, 4))[2] == 0    # cached
