"""
Serving form: a request loop that advects jobs of mixed sizes with continuous batching.

A job is (grid, number of steps, arrival time).  The read-only fields of a grid (V0, V1, Qx, Qy, ok) and its initial
tracer stack stay resident in HBM for the whole run; a job owns only its two state buffers and its workspace.  The
loop runs on one host thread:

* jobs whose arrival time has passed join the FIFO queue;
* queued jobs are admitted, in order, while the HBM budget and the slot count allow (a job that does not fit waits,
  and so do the ones behind it, so the order of service is the order of arrival);
* all admitted jobs advance together, one FCT step per batched launch (``batch.Batch``); when the set of admitted jobs
  changes the batch table is rebuilt, so a job that arrives while others are on the GPU joins at the next step;
* a job that has done its steps leaves: its response is its per-tracer content sums (the state stays on the device
  for the caller), its latency is the time from arrival to the end of its last step, and its memory returns to the
  budget.

Each job's result is bit-identical to the same job run alone (``advance_alone``): a batch member's arithmetic does
not depend on the other members.  In split precision each step is within 1e-5 of each tracer's range of the FP64
step; that bound holds per step, not over a job's whole run, in which any perturbation grows with the flow.
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import batch, ops, synth


@dataclass
class Grid:
    """A resident domain: read-only fields and the initial tracer stack, on the GPU."""
    V0: torch.Tensor
    V1: torch.Tensor
    Qx: torch.Tensor
    Qy: torch.Tensor
    ok: torch.Tensor
    C0: torch.Tensor
    dt: float
    flags: int = 0
    name: str = ""

    @classmethod
    def from_case(cls, case: synth.Case, device=None) -> "Grid":
        dev = device or ops.require_gpu()
        d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)     # noqa: E731
        return cls(V0=d(case.V0), V1=d(case.V1), Qx=d(case.Qx), Qy=d(case.Qy), ok=d(case.ok.astype(np.uint8)),
                   C0=d(case.C), dt=case.dt, name=case.info.get("name", ""))

    def resident_bytes(self) -> int:
        return sum(t.numel() * t.element_size() for t in (self.V0, self.V1, self.Qx, self.Qy, self.ok, self.C0))

    def job_bytes(self, precision: str) -> int:
        nt, nz, ny, nx = self.C0.shape
        L = ops.lib()
        ws = (L.ocx_fct_split_workspace_bytes if precision == "split" else L.ocx_fct_workspace_bytes)(nt, nz, ny, nx)
        return 2 * self.C0.numel() * 8 + int(ws)


@dataclass
class Job:
    id: int
    grid: int
    steps: int
    arrival: float                       # seconds after the start of the run
    admitted: float = math.nan
    done: float = math.nan
    admit_step: int = -1                 # index of the first batched step the job took part in
    done_step: int = -1                  # index of its last
    steps_done: int = 0
    content: np.ndarray | None = None    # response: per-tracer content sums V1 c at the end
    state: torch.Tensor | None = field(default=None, repr=False)
    alt: torch.Tensor | None = field(default=None, repr=False)
    ws: torch.Tensor | None = field(default=None, repr=False)
    bytes: int = 0

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def _regional_case(args):
    nx, ny, nz, nt, seed = args
    return synth.regional(nx=nx, ny=ny, nz=nz, nt=nt, seed=seed)


def make_grids(shapes, nt: int = 10, seed: int = 100, workers: int = 8, log=None) -> list[Grid]:
    """Resident grids for ``shapes`` [(nx, ny, nz)]: seeded regional domains (own mask, streamfunction and tracers
    each), generated on the host by ``workers`` threads of this process (NumPy releases the interpreter lock in its
    array work; no child process is started, so nothing is forked from a process that holds the GPU) and uploaded
    in order as they are done."""
    from concurrent.futures import ThreadPoolExecutor
    dev = ops.require_gpu()
    jobs = [(nx, ny, nz, nt, seed + m) for m, (nx, ny, nz) in enumerate(shapes)]
    out = []
    if workers <= 1:
        for c in map(_regional_case, jobs):
            out.append(Grid.from_case(c, dev))
        return out
    with ThreadPoolExecutor(max_workers=workers) as pool:
        for m, c in enumerate(pool.map(_regional_case, jobs)):
            out.append(Grid.from_case(c, dev))
            if log and (m + 1) % 16 == 0:
                log(f"grids {m + 1}/{len(jobs)}")
    return out


def poisson_jobs(n_jobs: int, rate: float, n_grids: int, steps_lo: int = 10, steps_hi: int = 500,
                 big_grid: int | None = None, big_prob: float = 0.02, seed: int = 7) -> list[Job]:
    """A seeded Poisson arrival stream: exponential gaps at ``rate`` jobs/s, step counts log-uniform in
    [steps_lo, steps_hi], a grid drawn uniformly (``big_grid`` with probability ``big_prob``)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_jobs))
    steps = np.exp(rng.uniform(np.log(steps_lo), np.log(steps_hi), n_jobs)).round().astype(int)
    small = [g for g in range(n_grids) if g != big_grid]
    grids = [int(big_grid) if big_grid is not None and rng.random() < big_prob else int(rng.choice(small))
             for _ in range(n_jobs)]
    return [Job(id=i, grid=grids[i], steps=int(steps[i]), arrival=float(t[i])) for i in range(n_jobs)]


class Server:
    """Continuous-batching request loop over resident grids."""

    def __init__(self, grids: list[Grid], hbm_budget: float = 200e9, max_active: int = 64, precision: str = "split",
                 rows_per_chunk: int = 0, waves_per_group: int = 0):
        if precision not in ops.PRECISIONS:
            raise ValueError(f"precision {precision!r}")
        self.grids = grids
        self.precision = precision
        self.max_active = max_active
        self.budget = float(hbm_budget) - sum(g.resident_bytes() for g in grids)
        if self.budget <= 0:
            raise ValueError("the resident grids alone exceed the HBM budget")
        self.tuning = (rows_per_chunk, waves_per_group)

    def _admit(self, job: Job) -> None:
        g = self.grids[job.grid]
        job.state = g.C0.clone()
        job.alt = torch.empty_like(g.C0)
        job.ws = torch.empty(max(g.job_bytes(self.precision) - 2 * g.C0.numel() * 8, 256), dtype=torch.uint8,
                             device=g.C0.device)
        job.bytes = g.job_bytes(self.precision)

    def _retire(self, job: Job) -> None:
        g = self.grids[job.grid]
        job.content = ops.content(job.state, g.V1).cpu().numpy()
        job.alt = None
        job.ws = None

    def _batch(self, active: list[Job]) -> batch.Batch:
        doms = []
        for j in active:
            g = self.grids[j.grid]
            doms.append(batch.Domain(C=j.state, V0=g.V0, V1=g.V1, Qx=g.Qx, Qy=g.Qy, ok=g.ok, dt=g.dt, flags=g.flags,
                                     precision=self.precision, C_alt=j.alt, workspace=j.ws))
        return batch.Batch(doms, rows_per_chunk=self.tuning[0], waves_per_group=self.tuning[1])

    def run(self, jobs: list[Job], keep_states: bool = False, log=None) -> dict:
        """Serve ``jobs`` (sorted by arrival); returns the statistics.  ``keep_states``: leave each job's final state
        on the device (``job.state``) for checks; otherwise it is dropped after the response."""
        pending = deque(sorted(jobs, key=lambda j: j.arrival))
        queue: deque[Job] = deque()
        active: list[Job] = []
        free = self.budget
        b = None
        steps_run = updates = 0
        t0 = time.perf_counter()
        busy = 0.0
        while pending or queue or active:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            changed = False
            while queue and len(active) < self.max_active:
                j = queue[0]
                need = self.grids[j.grid].job_bytes(self.precision)
                if need > free:
                    if not active:
                        raise RuntimeError(f"job {j.id} needs {need / 1e9:.1f} GB, more than the whole budget")
                    break
                queue.popleft()
                self._admit(j)
                j.admitted = time.perf_counter() - t0
                j.admit_step = steps_run
                free -= need
                active.append(j)
                changed = True
            if not active:
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - (time.perf_counter() - t0), 0.05)))
                continue
            if changed or b is None:
                if b is not None:
                    b.close()
                b = self._batch(active)
            ts = time.perf_counter()
            b.step()
            torch.cuda.current_stream().synchronize()
            busy += time.perf_counter() - ts
            steps_run += 1
            updates += b.updates()
            for m, j in enumerate(active):
                j.state, j.alt = b.state(m), b.spare(m)
                j.steps_done += 1
            finished = [j for j in active if j.steps_done >= j.steps]
            if finished:
                tdone = time.perf_counter() - t0
                for j in finished:
                    j.done = tdone
                    j.done_step = steps_run - 1
                    self._retire(j)
                    if not keep_states:
                        j.state = None
                    free += j.bytes
                active = [j for j in active if j.steps_done < j.steps]
                b.close()
                b = None
                if log:
                    log(f"t={tdone:7.2f}s done {len(finished)} active {len(active)} queued {len(queue)} "
                        f"pending {len(pending)}")
        wall = time.perf_counter() - t0
        if b is not None:
            b.close()
        lat = np.array([j.latency for j in jobs])
        return {"jobs": len(jobs), "wall_s": wall, "batched_steps": steps_run, "updates": int(updates),
                "updates_per_s": updates / wall, "gpu_busy_frac": busy / wall,
                "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
                "mean_latency_s": float(lat.mean()), "max_latency_s": float(lat.max())}


def advance_alone(grid: Grid, steps: int, precision: str) -> torch.Tensor:
    """The same job run alone through the single-domain entry point."""
    C = grid.C0.clone()
    out = torch.empty_like(C)
    ws = ops.Workspace()
    for _ in range(steps):
        ops.advect_fct(C, grid.V0, grid.V1, grid.Qx, grid.Qy, grid.ok, grid.dt, flags=grid.flags, out=out,
                       workspace=ws, precision=precision)
        C, out = out, C
    return C
