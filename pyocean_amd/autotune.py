"""
Launch-shape autotuner for the FCT step (single domain and batched).

The kernels march 64-lane wavefronts (63 cells of a row, one x-strip) up a chunk of rows of one layer; a workgroup
holds ``waves_per_group`` wavefronts, one tracer each, over the same strip, chunk and layer, so that they share the
per-cell fields (Qx, Qy, V0, V1, ok) through the L1/L2.  A launch shape is therefore

* ``rows_per_chunk`` - the y extent of one work item (each chunk re-reads a 3-row halo at its ends, so longer
  chunks move fewer bytes but give fewer work items to spread over 256 compute units);
* ``waves_per_group`` - the lane width of a workgroup, 64 x {1, 2, 4} lanes;
* ``groups_per_cu`` (batches only) - the persistent grid, in workgroups per compute unit.

The candidates are timed with HIP events on the caller's own workload (every launch shape gives bit-identical
results, so tuning changes speed only).  Results are cached in a JSON file (``OCX_TUNING_CACHE`` or
``~/.cache/pyocean_amd/tuning.json``) keyed by device name, precision and shape.
"""
from __future__ import annotations

import itertools
import json
import os
from dataclasses import asdict, dataclass
from pathlib import Path

import torch

from . import batch, ops

ROWS = (16, 32, 64, 128, 256)
WAVES = (1, 2, 4)
GROUPS = (1, 2, 4, 8)


@dataclass(frozen=True)
class Shape:
    rows_per_chunk: int
    waves_per_group: int
    groups_per_cu: int = 0

    @property
    def tile(self) -> str:
        """The work item as (x lanes) x (y rows) x (layers)."""
        return f"{64 * self.waves_per_group}x{self.rows_per_chunk}x1"


def cache_path() -> Path:
    p = os.environ.get("OCX_TUNING_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "pyocean_amd" / "tuning.json"


def _load() -> dict:
    p = cache_path()
    try:
        return json.loads(p.read_text())
    except (OSError, ValueError):
        return {}


def _store(key: str, shape: Shape) -> None:
    p = cache_path()
    data = _load()
    data[key] = asdict(shape)
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        tmp = p.with_suffix(".tmp")
        tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
        tmp.replace(p)
    except OSError:
        pass                                   # a read-only home: tuning still applies to this process


def _key(kind: str, precision: str, shapes) -> str:
    dev = torch.cuda.get_device_name(torch.cuda.current_device()).replace(" ", "_")
    return f"{dev}|{kind}|{precision}|" + ",".join("x".join(str(v) for v in s) for s in shapes)


def cached(kind: str, precision: str, shapes) -> Shape | None:
    v = _load().get(_key(kind, precision, shapes))
    return Shape(**v) if v else None


def _time_ms(fn, reps: int, stream) -> float:
    """Median of ``reps`` timed calls (after one untimed call), HIP events on ``stream``."""
    fn()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for a, b in ev:
        a.record(stream)
        fn()
        b.record(stream)
    stream.synchronize()
    t = sorted(a.elapsed_time(b) for a, b in ev)
    return t[len(t) // 2]


def tune_step(C, V0, V1, Qx, Qy, ok, dt, precision: str = "f64", flags: int = 0, inflow=None, reps: int = 3,
              rows=ROWS, waves=WAVES, use_cache: bool = True, log=None):
    """Tune the single-domain step (``ops.advect_fct``) on these inputs; leaves the best shape set
    (``batch.set_tuning``) and returns (best, table) with table = [(shape, ms)].  C is not modified."""
    shapes = [tuple(C.shape)]
    if use_cache:
        hit = cached("step", precision, shapes)
        if hit is not None:
            batch.set_tuning(hit.rows_per_chunk, hit.waves_per_group)
            return hit, []
    out = torch.empty_like(C)
    ws = ops.Workspace()
    stream = torch.cuda.current_stream()
    table = []
    for r, w in itertools.product(rows, waves):
        batch.set_tuning(r, w)
        ms = _time_ms(lambda: ops.advect_fct(C, V0, V1, Qx, Qy, ok, dt, inflow=inflow, flags=flags, out=out,
                                             workspace=ws, precision=precision), reps, stream)
        table.append((Shape(r, w), ms))
        if log:
            log(f"tune step {Shape(r, w).tile}: {ms:.3f} ms")
    best = min(table, key=lambda t: t[1])[0]
    batch.set_tuning(best.rows_per_chunk, best.waves_per_group)
    _store(_key("step", precision, shapes), best)
    return best, table


def tune_batch(domains: list[batch.Domain], reps: int = 2, rows=ROWS, waves=WAVES, groups=GROUPS,
               graph: bool = False, use_cache: bool = True, log=None):
    """Tune the batched step over ``domains``.  Returns (best, table, steps_taken): every trial step advances the
    domains' states (``steps_taken`` steps in all; each launch shape computes the same states, so the domains end
    exactly as after that many untuned steps, in ``domains[d].C`` since the count is even)."""
    shapes = [tuple(d.C.shape) + (d.precision == "split",) for d in domains]
    prec = "mixed" if len({d.precision for d in domains}) > 1 else domains[0].precision
    if use_cache:
        hit = cached("batch", prec, shapes)
        if hit is not None:
            return hit, [], 0
    stream = torch.cuda.current_stream()
    if graph and stream.cuda_stream == 0:
        stream = torch.cuda.Stream()                  # graph capture needs a created stream
        stream.wait_stream(torch.cuda.current_stream())
    table = []
    steps = 0
    # the second state buffers and workspaces are allocated once and shared by every candidate batch
    doms = list(domains)
    probe = batch.Batch(doms)
    for m, d in enumerate(doms):
        d.C_alt = probe.spare(m)
        d.workspace = probe.workspace(m)
    probe.close()
    for r, w, g in itertools.product(rows, waves, groups):
        b = batch.Batch(doms, rows_per_chunk=r, waves_per_group=w, groups_per_cu=g)

        def two():
            b.step(graph=graph, stream=stream)
            b.step(graph=graph, stream=stream)
        ms = _time_ms(two, reps, stream) / 2
        steps += 2 * (reps + 1)
        b.close()
        shape = Shape(r, w, g)
        table.append((shape, ms))
        if log:
            log(f"tune batch {shape.tile} groups/CU {g}: {ms:.2f} ms")
    stream.synchronize()
    best = min(table, key=lambda t: t[1])[0]
    _store(_key("batch", prec, shapes), best)
    return best, table, steps
