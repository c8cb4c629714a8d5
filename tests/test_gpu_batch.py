"""
GPU tests of the batched form: every member of a batch (uneven sizes, FP64 and split members, open and closed
edges, inflow) ends bit-identical to the same domain advanced alone, eagerly and through the captured graph, over
several steps with other device work between graph replays.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from pyocean_amd import batch, ops, synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu case]", *a, flush=True)


def _domain(nx, ny, nz, nt, seed, precision, open_mask=0, inflow=False):
    c = synth.regional(nx=nx, ny=ny, nz=nz, nt=nt, seed=seed)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()        # noqa: E731
    Qx, Qy = c.Qx.copy(), c.Qy.copy()
    V1 = c.V1
    if open_mask:
        # open the west/south edges with inflow through them; V1 from the same continuity
        Qx[..., 0] = 0.02 * np.abs(c.Qx).max()
        Qy[:, 0, :] = 0.02 * np.abs(c.Qy).max()
        Qx[..., 0] *= c.V0[..., 0] > 0
        Qy[:, 0, :] *= c.V0[:, 0, :] > 0
        V1 = c.V0 - c.dt * (Qx[..., 1:] - Qx[..., :-1] + Qy[:, 1:] - Qy[:, :-1])
    ok = (c.V0 > 0) & (V1 > 0)
    return batch.Domain(C=d(c.C), V0=d(c.V0), V1=d(V1), Qx=d(Qx), Qy=d(Qy), ok=d(ok.astype(np.uint8)), dt=c.dt,
                        flags=open_mask, precision=precision, inflow=d(c.C * 0.9 + 0.1) if inflow else None)


def _solo(dom, steps):
    C = dom.C.clone()
    for _ in range(steps):
        C = ops.advect_fct(C, dom.V0, dom.V1, dom.Qx, dom.Qy, dom.ok, dom.dt, inflow=dom.inflow, flags=dom.flags,
                           precision=dom.precision)
    return C


@pytest.mark.parametrize("graph", [False, True])
def test_batch_members_equal_their_solo_runs(gpu, graph):
    _say("batch", "graph" if graph else "eager")
    shapes = [(130, 97, 7, 3, "f64", 0, False), (64, 200, 5, 2, "split", 5, True), (201, 66, 9, 4, "f64", 5, False),
              (96, 96, 4, 1, "split", 0, False), (70, 150, 6, 5, "f64", 0, True)]
    doms = [_domain(nx, ny, nz, nt, 40 + m, pr, om, inf) for m, (nx, ny, nz, nt, pr, om, inf) in enumerate(shapes)]
    solo0 = [d.C.clone() for d in doms]
    b = batch.Batch(doms, rows_per_chunk=32)
    stream = torch.cuda.Stream() if graph else torch.cuda.current_stream()
    steps = 4
    with torch.cuda.stream(stream):
        for k in range(steps):
            b.step(graph=graph, stream=stream)
            stream.synchronize()
            if graph:                                    # eager device work between the replays
                torch.cuda.current_stream().synchronize()
                _ = (doms[0].C * 2.0).sum().item()
    for m, d in enumerate(doms):
        d0 = batch.Domain(C=solo0[m], V0=d.V0, V1=d.V1, Qx=d.Qx, Qy=d.Qy, ok=d.ok, dt=d.dt, flags=d.flags,
                          precision=d.precision, inflow=d.inflow)
        ref = _solo(d0, steps)
        assert torch.equal(b.state(m), ref), (m, shapes[m])
    b.close()


def test_batch_graph_replays_match_eager_each_step(gpu):
    _say("batch graph vs eager per step")
    mk = lambda: [_domain(150, 120, 6, 3, 70, "f64"), _domain(90, 130, 8, 2, 71, "split", 5, True)]   # noqa: E731
    eager, graphed = batch.Batch(mk()), batch.Batch(mk())
    s = torch.cuda.Stream()
    for k in range(5):
        eager.step()
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            graphed.step(graph=True, stream=s)
        s.synchronize()
        for m in range(2):
            assert torch.equal(eager.state(m), graphed.state(m)), (k, m)
    eager.close()
    graphed.close()
