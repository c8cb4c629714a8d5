"""
GPU tests of the batched form: every member of a batch (uneven sizes, FP64 and split members, open and closed
edges, inflow) ends bit-identical to the same domain advanced alone, eagerly and through the captured graph, over
several steps with other device work between graph replays; FP64 members equal the oracle bit for bit and split
members stay within 1e-5 of each tracer's span of it, conservative and bounded; and the full 64-domain mixed batch
(nx, ny log-uniform in [96, 768], nz in {20, 30, 50, 75}, ten tracers, 3.24 G tracer-cells) conserves and bounds
every member.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from pyocean_amd import batch, ops, synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu case]", *a, flush=True)


def _host(nx, ny, nz, nt, seed, open_mask=0, inflow=False):
    """The member's host arrays (also the oracle's inputs)."""
    c = synth.regional(nx=nx, ny=ny, nz=nz, nt=nt, seed=seed)
    Qx, Qy = c.Qx.copy(), c.Qy.copy()
    V1 = c.V1
    if open_mask:
        # open the west/south edges with inflow through them, each layer's share in proportion to its edge cell's
        # volume (so the 1 mm sub-bed layers take a 1 mm share, as the open-edge layer transports do); V1 from the
        # same continuity
        Qx[..., 0] = 0.02 * np.abs(c.Qx).max() * c.V0[..., 0] / c.V0[..., 0].max()
        Qy[:, 0, :] = 0.02 * np.abs(c.Qy).max() * c.V0[:, 0, :] / c.V0[:, 0, :].max()
        V1 = c.V0 - c.dt * (Qx[..., 1:] - Qx[..., :-1] + Qy[:, 1:] - Qy[:, :-1])
    ok = (c.V0 > 0) & (V1 > 0)
    return dict(C=c.C, V0=c.V0, V1=V1, Qx=Qx, Qy=Qy, ok=ok, dt=c.dt, open_mask=open_mask,
                inflow=c.C * 0.9 + 0.1 if inflow else None)


def _device(h, precision):
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()        # noqa: E731
    return batch.Domain(C=d(h["C"]), V0=d(h["V0"]), V1=d(h["V1"]), Qx=d(h["Qx"]), Qy=d(h["Qy"]),
                        ok=d(h["ok"].astype(np.uint8)), dt=h["dt"], flags=h["open_mask"], precision=precision,
                        inflow=None if h["inflow"] is None else d(h["inflow"]))


def _domain(nx, ny, nz, nt, seed, precision, open_mask=0, inflow=False):
    return _device(_host(nx, ny, nz, nt, seed, open_mask, inflow), precision)


SHAPES = [(130, 97, 7, 3, "f64", 0, False), (64, 200, 5, 2, "split", 5, True), (201, 66, 9, 4, "f64", 5, False),
          (96, 96, 4, 1, "split", 0, False), (70, 150, 6, 5, "f64", 0, True)]


def _solo(dom, steps):
    C = dom.C.clone()
    for _ in range(steps):
        C = ops.advect_fct(C, dom.V0, dom.V1, dom.Qx, dom.Qy, dom.ok, dom.dt, inflow=dom.inflow, flags=dom.flags,
                           precision=dom.precision)
    return C


@pytest.mark.parametrize("graph", [False, True])
def test_batch_members_equal_their_solo_runs(gpu, graph):
    _say("batch", "graph" if graph else "eager")
    shapes = SHAPES
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


def test_batch_members_match_the_oracle(gpu, oracle_c):
    # one batched step (captured graph) of the five mixed members against the C oracle's FP64 step: FP64 members
    # bit for bit, split members within 1e-5 of each tracer's span, with closed-edge content kept to 1e-13 and no
    # new extrema (bounds over the member's values and its inflow)
    _say("batch vs oracle")
    hosts = [_host(nx, ny, nz, nt, 40 + m, om, inf) for m, (nx, ny, nz, nt, _, om, inf) in enumerate(SHAPES)]
    doms = [_device(h, sh[4]) for h, sh in zip(hosts, SHAPES)]
    b = batch.Batch(doms, rows_per_chunk=32)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        b.step(graph=True, stream=s)
    s.synchronize()
    for m, (h, sh) in enumerate(zip(hosts, SHAPES)):
        ref = oracle_c.advect_fct(h["C"], h["V0"], h["V1"], h["Qx"], h["Qy"], h["ok"], h["dt"], inflow=h["inflow"],
                                  open_mask=h["open_mask"])
        got = b.state(m).cpu().numpy()
        if sh[4] == "f64":
            assert np.array_equal(got, ref), (m, sh)
            continue
        okb = np.broadcast_to(h["ok"], h["C"].shape)
        for t in range(h["C"].shape[0]):
            vals = h["C"][t][okb[t]]
            if h["inflow"] is not None:
                vals = np.concatenate([vals, h["inflow"][t].ravel()])
            lo, hi = vals.min(), vals.max()
            span = hi - lo
            assert np.abs(got[t] - ref[t])[okb[t]].max() <= 1e-5 * span, (m, t)
            assert got[t][okb[t]].min() >= lo - 1e-12 * span and got[t][okb[t]].max() <= hi + 1e-12 * span
            if not h["open_mask"]:
                c0 = (h["V0"] * h["C"][t]).sum()
                assert abs((h["V1"] * got[t]).sum() - c0) <= 1e-13 * abs(c0)
    b.close()


def test_batch_member_flags_outside_the_batch_step_are_refused(gpu):
    # a batch runs the limited FCT step: a member asking for the unlimited step, or any flag bit beyond the open
    # edges and WENO5-JS, is refused on the host before any launch
    _say("batch flags")
    dom = _domain(70, 60, 3, 1, 5, "f64")
    for bad in (ops.flags_of(0, unlimited=True), 1 << 6, 1 << 31):
        dom.flags = bad
        with pytest.raises(ValueError, match="flags"):
            batch.Batch([dom])
    dom.flags = ops.flags_of(15, js=True)
    b = batch.Batch([dom])
    b.close()


def test_full_64_domain_mixed_batch_conserves_and_bounds_every_member(gpu):
    # the batched configuration itself: 64 regional domains (synth.batch_shapes(64, seed=7), seeds 100 + d, ten
    # tracers), even members FP64 and odd members split, one captured-graph step; every member keeps its content to
    # 1e-13 per tracer and makes no new extrema, and three members equal their solo runs bit for bit
    from pyocean_amd import serve
    shapes = synth.batch_shapes(64, seed=7)
    _say("full 64-domain mixed batch: generating")
    grids = serve.make_grids(shapes, nt=10, seed=100, workers=16)
    prec = ["f64" if d % 2 == 0 else "split" for d in range(len(grids))]
    doms = [batch.Domain(C=g.C0, V0=g.V0, V1=g.V1, Qx=g.Qx, Qy=g.Qy, ok=g.ok, dt=g.dt, precision=p)
            for g, p in zip(grids, prec)]
    assert sum(d.cells for d in doms) > 3.2e9
    _say("full 64-domain mixed batch: one step")
    b = batch.Batch(doms)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        b.step(graph=True, stream=s)
    s.synchronize()
    for m, (g, d) in enumerate(zip(grids, doms)):
        out = b.state(m)
        m0 = ops.content(g.C0, g.V0).cpu().numpy()
        m1 = ops.content(out, g.V1).cpu().numpy()
        np.testing.assert_allclose(m1, m0, rtol=1e-13, err_msg=f"domain {m} {shapes[m]} {prec[m]}")
        okb = torch.broadcast_to(g.ok.bool(), g.C0.shape)
        for t in range(g.C0.shape[0]):
            c0, c1 = g.C0[t][okb[t]], out[t][okb[t]]
            lo, hi = c0.min().item(), c0.max().item()
            span = hi - lo
            assert c1.min().item() >= lo - 1e-12 * span and c1.max().item() <= hi + 1e-12 * span, (m, t)
    big = max(range(len(doms)), key=lambda m: doms[m].cells)
    for m in sorted({0, 1, big}):
        d = doms[m]
        solo = ops.advect_fct(d.C, d.V0, d.V1, d.Qx, d.Qy, d.ok, d.dt, precision=d.precision)
        assert torch.equal(b.state(m), solo), (m, shapes[m], prec[m])
    b.close()
