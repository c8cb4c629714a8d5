"""
Benchmark of the tracer-advection step on one MI355X.

Workload: the regional case, 1024 x 1024 x 50 curvilinear grid, ~20 % land, z* layers, ten tracers
[T, S, 8 passive], plain FP64 WENO5-Z + Zalesak FCT + SSP-RK3 (``ocx_advect_fct_f64``), inputs resident in HBM.
One step = one FCT step of the whole stack (C_{n+1} = step(C_n)); the metric counts every grid cell, land
included: nt * nz * ny * nx tracer-cell updates per step.

Prints ONE JSON line with the throughput, per-step p50/p99 latency, the roofline of the dominant kernel (per-kernel
durations from hipEvents recorded by the library around each launch on its stream), the step-level roofline with
the algorithmic bytes per update B = 72 + 99/nt, and a CPU baseline (the oracle's C loops on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12                 # MI355X HBM3E, bytes/s
METRIC = "tracer-cell updates/s (1024²×50, 10 tracers); % HBM roofline; p50/p99"


def step_bytes(nt: int, precision: str = "f64") -> float:
    """Algorithmic HBM bytes per tracer-cell update of one FCT step (three dependent RK3 sweeps, FCT fused into the
    third): tracer streams 48 B read + 24 B written, shared Qx, Qy, V0, V1 (32 B) and ok (1 B) read once per sweep;
    with c1 and c2 stored in FP32, 52 B instead of 72."""
    return (52.0 if precision == "split" else 72.0) + 99.0 / nt


# compulsory bytes per tracer-cell of each kernel of the step (every array it touches moved once):
# (bytes per tracer-cell, shared bytes per cell amortised over the tracers)
KERNEL_BYTES = {
    "f64": {
        "fct_stage1": (32.0, 33.0),        # C -> c1, F0 (x, y); Qx, Qy, V0, V1, ok
        "fct_stage2": (56.0, 33.0),        # c1, C, F0 -> c2, S = F0 + F1
        "fct_stage3_split": (56.0, 33.0),  # c2, S, C -> A (x, y), c_L
        "fct_limit_final": (40.0, 33.0),   # C, c_L, A -> out; Qx, Qy, V0, V1, ok
    },
    "split": {                             # c1, c2, F0, S, A in FP32; C, c_L, out in FP64
        "fct_stage1": (20.0, 33.0),
        "fct_stage2": (32.0, 33.0),
        "fct_stage3_split": (36.0, 33.0),
        "fct_limit_final": (32.0, 33.0),
    },
}


def read_traffic(kernel: str, workload: str):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3 --pmc summary, scaled by the copy-kernel
    calibration written in the same file; None when no summary for this workload exists."""
    path = ROOT / "profiles" / f"traffic_{workload}.txt"
    if not path.exists():
        return None
    sf, sw, fetch, write = 1.0, 1.0, None, None
    for line in path.read_text().splitlines():
        parts = line.split()
        if len(parts) >= 2 and parts[0] == "scale_fetch":
            sf = float(parts[1])
        if len(parts) >= 2 and parts[0] == "scale_write":
            sw = float(parts[1])
        if len(parts) >= 3 and parts[0] == kernel:
            fetch, write = float(parts[1]), float(parts[2])
    if fetch is None:
        return None
    return (fetch * sf + write * sw) * 1024.0


def cpu_model() -> dict:
    """The host CPU: model name (lscpu / /proc/cpuinfo), logical CPUs, and the CPUs this process may run on."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def _time_fct(cref, C, case, nz, reps):
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        cref.advect_fct(C, case.V0[:nz], case.V1[:nz], case.Qx[:nz], case.Qy[:nz], case.ok[:nz], case.dt)
        times.append(time.perf_counter() - t0)
    return float(np.median(times))


def cpu_baseline(case, nt_sample: int, reps: int):
    """The oracle's C loops (OpenMP over (tracer, layer) planes, like the reference's prange) on bounded samples of
    the same workload: the first ``nt_sample`` tracers with all threads the process may use (median of ``reps``
    steps), and one tracer of the first 10 layers on one thread (one step).  Tracers and layers are independent
    work items of the loops, so both rates scale to the whole stack."""
    from oracle import cref
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    cref.set_threads(threads)
    C = np.ascontiguousarray(case.C[:nt_sample])
    nz = case.C.shape[1]
    t = _time_fct(cref, C, case, nz, reps)
    n = C.size
    cref.set_threads(1)
    nz1 = min(10, nz)
    C1 = np.ascontiguousarray(case.C[:1, :nz1])
    t1 = _time_fct(cref, C1, case, nz1, 1)
    cref.set_threads(threads)
    return {"value": n / t, "unit": "tracer-cell updates/s", "cores": threads, "kind": "port",
            "sample": f"{nt_sample} of {case.C.shape[0]} tracers of the same {case.info['name']} inputs, one FCT "
                      f"step, median of {reps} ({t:.2f} s each); C loops of the oracle, -O3 -ffp-contract=off, "
                      f"OpenMP over (tracer, layer) planes",
            "one_thread": {"value": C1.size / t1, "unit": "tracer-cell updates/s", "cores": 1,
                           "sample": f"1 tracer x the first {nz1} layers, one FCT step ({t1:.2f} s)"},
            "cpu": cpu_model()}


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def bench_batch(args) -> int:
    """64 regional domains (nx, ny log-uniform in [96, 768], nz in {20, 30, 50, 75}, ten tracers, seeds 100 + d),
    FP64 and split members, advanced by the batched step as a captured hipGraph."""
    import torch

    from pyocean_amd import _native, autotune, batch, ops, serve, synth

    shapes = synth.batch_shapes(args.domains, seed=7) if not args.same_shape else [(args.nx, args.ny, args.nz)] * args.domains
    _log(f"generating {len(shapes)} domains")
    grids = serve.make_grids(shapes, nt=args.nt, seed=100, workers=args.workers, log=_log)
    prec = [("f64" if d % 2 == 0 else "split") if args.mix == "mixed" else args.mix for d in range(len(grids))]
    doms = [batch.Domain(C=g.C0, V0=g.V0, V1=g.V1, Qx=g.Qx, Qy=g.Qy, ok=g.ok, dt=g.dt, precision=p)
            for g, p in zip(grids, prec)]
    updates = sum(d.cells for d in doms)
    shape = autotune.Shape(args.rch, args.wpb, 0)
    tuned = None
    if args.autotune:
        _log("autotuning the launch shape")
        shape, table, _ = autotune.tune_batch(doms, reps=1, graph=not args.eager, use_cache=False, log=_log)
        tuned = {"best": shape.tile, "groups_per_cu": shape.groups_per_cu,
                 "table_ms": {f"{s.tile}/g{s.groups_per_cu}": round(ms, 3) for s, ms in table}}
    b = batch.Batch(doms, rows_per_chunk=shape.rows_per_chunk, waves_per_group=shape.waves_per_group,
                    groups_per_cu=shape.groups_per_cu)
    stream = torch.cuda.Stream()                      # graph capture needs a created stream
    torch.cuda.set_stream(stream)
    graph = not args.eager
    _log(f"{updates / 1e9:.2f} G tracer-cells per step; warmup {args.warmup}")
    for _ in range(max(args.warmup, 1)):
        b.step(graph=graph, stream=stream)
    # per-kernel durations from two eager steps
    torch.cuda.synchronize()
    _native.profile_enable(True)
    _native.profile_collect()
    for _ in range(2):
        b.step(stream=stream)
    torch.cuda.synchronize()
    _native.profile_enable(False)
    prof = _native.profile_collect()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        b.step(graph=graph, stream=stream)
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = np.array([a.elapsed_time(c) for a, c in ev])
    value = updates * args.steps / wall
    names = {"batch_stage1": "fct_stage1", "batch_stage2": "fct_stage2", "batch_stage3_split": "fct_stage3_split",
             "batch_limit_final": "fct_limit_final"}
    kernels = {k: v[0] / v[1] for k, v in prof.items()}
    dom = max(kernels, key=kernels.get)

    def kbytes(kname):
        tot = 0.0
        for d in doms:
            per_tc, shared = KERNEL_BYTES[d.precision][names.get(kname, "fct_stage1")]
            tot += d.cells * (per_tc + shared / d.C.shape[0])
        return tot

    alg = kbytes(dom)
    achieved = alg / (kernels[dom] * 1e-3)
    B = sum(d.cells * step_bytes(d.C.shape[0], d.precision) for d in doms) / updates
    b.close()
    res = {
        "metric": "tracer-cell updates/s (batched: 64 regional domains of uneven size, one persistent launch)",
        "value": value, "unit": "tracer-cell updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "p50_ms": float(np.percentile(step_ms, 50)),
        "p99_ms": float(np.percentile(step_ms, 99)), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": {"mixed": "f64 + f32/f64", "split": "f32/f64", "f64": "f64"}[args.mix],
        "data": "synthetic (seeded regional domains, seeds 100 + d)",
        "config": {"workload": f"{len(doms)} regional domains, nx, ny log-uniform in [96, 768], nz in "
                               f"{{20, 30, 50, 75}}, {args.nt} tracers, members {args.mix}, "
                               + ("eager launches" if args.eager else "hipGraph step"),
                   "tracer_cells": updates, "rows_per_chunk": shape.rows_per_chunk,
                   "waves_per_group": shape.waves_per_group, "groups_per_cu": shape.groups_per_cu},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None, "alg_bytes_per_launch": alg,
                     "avg_launch_ms": kernels[dom]},
        "step_roofline": {"bytes_per_update": B, "achieved": value * B / 1e9, "peak": HBM_PEAK / 1e9,
                          "unit": "GB/s", "frac": value * B / HBM_PEAK},
        "kernels_ms": kernels,
        "autotune": tuned,
        "cpu_baseline": None,
    }
    print(json.dumps(res), flush=True)
    return 0


def bench_serve(args) -> int:
    """The serving stream: Poisson arrivals at ``--rate`` jobs/s, grids as in the batch form plus a 1536 x 1536 x 50
    domain drawn with probability 0.02, steps log-uniform in [10, 500], split precision, HBM budget 200 GB."""
    from pyocean_amd import serve, synth

    shapes = synth.batch_shapes(args.domains, seed=7) + [(1536, 1536, 50)]
    _log(f"generating {len(shapes)} grids")
    grids = serve.make_grids(shapes, nt=args.nt, seed=100, workers=args.workers, log=_log)
    jobs = serve.poisson_jobs(args.jobs, args.rate, len(grids), steps_lo=10, steps_hi=500, big_grid=len(grids) - 1,
                              big_prob=0.02, seed=7)
    srv = serve.Server(grids, hbm_budget=args.budget_gb * 1e9, max_active=64, precision="split",
                       rows_per_chunk=args.rch, waves_per_group=args.wpb)
    _log(f"serving {len(jobs)} jobs at {args.rate} jobs/s")
    st = srv.run(jobs, log=_log)
    res = {
        "metric": "serving: tracer-cell updates/s and per-job latency p50/p99 (Poisson arrivals, continuous batching)",
        "value": st["updates_per_s"], "unit": "tracer-cell updates/s", "n_gpus": 1, "steps": st["batched_steps"],
        "warmup": 0, "ms_per_step": st["wall_s"] / max(st["batched_steps"], 1) * 1e3,
        "p50_latency_s": st["p50_latency_s"], "p99_latency_s": st["p99_latency_s"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32/f64",
        "data": "synthetic (seeded regional domains, seeded Poisson stream)",
        "config": {"workload": f"{len(jobs)} jobs, Poisson {args.rate}/s, {len(grids) - 1} regional grids + "
                               "1536x1536x50 (p = 0.02), steps log-uniform in [10, 500], split precision, "
                               f"HBM budget {args.budget_gb:.0f} GB, at most 64 jobs in a batch",
                   "precision_setting": "split: within 1e-5 of each tracer's range of the FP64 oracle per step "
                                        "(tested); over a job's many steps the difference grows as the flow "
                                        "amplifies any perturbation (bench.py --form precision)"},
        "serve": st,
        "cpu_baseline": None,
    }
    print(json.dumps(res), flush=True)
    return 0


def bench_precision(args) -> int:
    """The split form against the plain FP64 form (bit-identical to the oracle) on the same inputs, stepped side by
    side: per-tracer max and RMS error relative to the tracer's initial range, content drift and global bounds."""
    import torch

    from pyocean_amd import ops, synth

    marks = sorted(int(v) for v in args.checkpoints.split(","))
    _log(f"generating regional {args.nx}x{args.ny}x{args.nz} nt={args.nt}")
    case = synth.regional(nx=args.nx, ny=args.ny, nz=args.nz, nt=args.nt, seed=0)
    dev = torch.device("cuda", 0)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)     # noqa: E731
    V0, V1, Qx, Qy, ok = d(case.V0), d(case.V1), d(case.Qx), d(case.Qy), d(case.ok.astype(np.uint8))
    wet = ok.bool()
    C0 = d(case.C)
    ref, ref_o = C0.clone(), torch.empty_like(C0)
    spl, spl_o = C0.clone(), torch.empty_like(C0)
    # control: the FP64 form from the initial state rounded to FP32 (a perturbation of the size of one FP32
    # rounding, about what one split step introduces), to separate the problem's own sensitivity to small
    # perturbations from the split form's arithmetic
    ctl, ctl_o = C0.float().double(), torch.empty_like(C0)
    ws_r, ws_s, ws_c = ops.Workspace(), ops.Workspace(), ops.Workspace()
    nt = C0.shape[0]
    w = wet.unsqueeze(0).expand_as(C0)
    lo = torch.where(w, C0, torch.full_like(C0, np.inf)).amin(dim=(1, 2, 3))
    hi = torch.where(w, C0, torch.full_like(C0, -np.inf)).amax(dim=(1, 2, 3))
    span = (hi - lo).cpu().numpy()
    m0 = ops.content(C0, V0).cpu().numpy()
    out = {}
    k = 0
    t_ref = t_spl = 0.0
    for mark in marks:
        while k < mark:
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            c = torch.cuda.Event(enable_timing=True)
            a.record()
            ops.advect_fct(ref, V0, V1, Qx, Qy, ok, case.dt, out=ref_o, workspace=ws_r, precision="f64")
            b.record()
            ops.advect_fct(spl, V0, V1, Qx, Qy, ok, case.dt, out=spl_o, workspace=ws_s, precision="split")
            c.record()
            ops.advect_fct(ctl, V0, V1, Qx, Qy, ok, case.dt, out=ctl_o, workspace=ws_c, precision="f64")
            ref, ref_o = ref_o, ref
            spl, spl_o = spl_o, spl
            ctl, ctl_o = ctl_o, ctl
            k += 1
            if k % 100 == 0:
                torch.cuda.synchronize()
                t_ref += a.elapsed_time(b)
                t_spl += b.elapsed_time(c)
                _log(f"step {k}")
        torch.cuda.synchronize()
        diff = torch.where(w, (spl - ref).abs(), torch.zeros_like(ref))
        nwet = w.sum(dim=(1, 2, 3)).double()
        emax = diff.amax(dim=(1, 2, 3)).cpu().numpy()
        erms = torch.sqrt((diff * diff).sum(dim=(1, 2, 3)) / nwet).cpu().numpy()
        # where the largest split error sits: layer, row, column and that cell's volume against the median
        tw = int(np.argmax(emax / span))
        flat = int(torch.argmax(diff[tw]).item())
        kk, jj, ii = np.unravel_index(flat, diff.shape[1:])
        v1med = float(V1[wet].median().item())
        where = {"tracer": tw, "k": int(kk), "j": int(jj), "i": int(ii),
                 "V1_over_median": float(V1[kk, jj, ii].item()) / v1med,
                 "V0_over_median": float(V0[kk, jj, ii].item()) / v1med}
        dctl = torch.where(w, (ctl - ref).abs(), torch.zeros_like(ref))
        cmax_ = dctl.amax(dim=(1, 2, 3)).cpu().numpy()
        crms = torch.sqrt((dctl * dctl).sum(dim=(1, 2, 3)) / nwet).cpu().numpy()
        del dctl
        # every step maps (V0, c^n) to (V1, c^{n+1}) with the same volumes, so the conserved quantity of the last
        # step is sum V0 c^{n-1} -> sum V1 c^n (spl_o still holds c^{n-1})
        ms = ops.content(spl, V1).cpu().numpy()
        ms_prev = ops.content(spl_o, V0).cpu().numpy()
        mr = ops.content(ref, V1).cpu().numpy()
        smin = torch.where(w, spl, torch.full_like(spl, np.inf)).amin(dim=(1, 2, 3)).cpu().numpy()
        smax = torch.where(w, spl, torch.full_like(spl, -np.inf)).amax(dim=(1, 2, 3)).cpu().numpy()
        lo_n, hi_n = lo.cpu().numpy(), hi.cpu().numpy()
        out[str(mark)] = {
            "max_rel_err": (emax / span).tolist(),
            "rms_rel_err": (erms / span).tolist(),
            "largest_split_error_at": where,
            "control_f64_from_f32_rounded_state_max_rel": (cmax_ / span).tolist(),
            "control_f64_from_f32_rounded_state_rms_rel": (crms / span).tolist(),
            "split_last_step_content_rel": (np.abs(ms - ms_prev) / np.abs(ms_prev)).tolist(),
            "split_vs_f64_content_rel": (np.abs(ms - mr) / np.abs(mr)).tolist(),
            "split_below_initial_min": (np.maximum(lo_n - smin, 0) / span).tolist(),
            "split_above_initial_max": (np.maximum(smax - hi_n, 0) / span).tolist(),
        }
        del diff
        _log(f"after {mark} steps: max rel err {np.max(emax / span):.3e}, rms {np.max(erms / span):.3e}; control "
             f"max {np.max(cmax_ / span):.3e}, rms {np.max(crms / span):.3e}")
    res = {
        "metric": "split vs FP64: per-tracer max and RMS error relative to the tracer's initial range",
        "value": float(max(max(v["max_rel_err"]) for v in out.values())), "unit": "relative error",
        "n_gpus": 1, "steps": marks[-1], "warmup": 0, "higher_is_better": False,
        "dtype": "f32/f64 against f64",
        "data": "synthetic (seeded regional case)",
        "config": {"workload": f"regional {args.nx}x{args.ny}x{args.nz}, {nt} tracers [T, S, 8 passive], closed "
                               "edges, the same (V0, V1, Qx, Qy) every step",
                   "tracers": ["T", "S"] + [f"tr{m}" for m in range(nt - 2)],
                   "initial_span": span.tolist(), "first_step_content": m0.tolist()},
        "checkpoints": out,
        "ms_per_step_sampled": {"f64": t_ref / max(k // 100, 1), "split": t_spl / max(k // 100, 1)},
    }
    print(json.dumps(res), flush=True)
    return 0


def bench_precision_solid(args) -> int:
    """The split form against the oracle itself (its C loops, stepped on the host) on the solid-body case: the
    128 x 128 x 20 closed basin with a Gaussian blob and a slotted cylinder in solid-body rotation at Courant 0.4,
    100 steps (one sixth of a revolution); per-tracer max and RMS error relative to the tracer's initial range at
    each checkpoint, content and bounds."""
    import torch

    from oracle import cref
    from pyocean_amd import ops, synth

    marks = sorted(int(v) for v in args.checkpoints.split(","))
    case = synth.solid_body()
    dev = torch.device("cuda", 0)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)     # noqa: E731
    V0, V1, Qx, Qy, ok = d(case.V0), d(case.V1), d(case.Qx), d(case.Qy), d(case.ok.astype(np.uint8))
    okb = np.broadcast_to(case.ok, case.C.shape)
    lo = np.array([case.C[t][okb[t]].min() for t in range(case.C.shape[0])])
    hi = np.array([case.C[t][okb[t]].max() for t in range(case.C.shape[0])])
    span = hi - lo
    m0 = (case.V0 * case.C).sum(axis=(1, 2, 3))
    spl = d(case.C)
    ref = case.C.copy()
    out = {}
    k = 0
    for mark in marks:
        while k < mark:
            spl = ops.advect_fct(spl, V0, V1, Qx, Qy, ok, case.dt, precision="split")
            ref = cref.advect_fct(ref, case.V0, case.V1, case.Qx, case.Qy, case.ok, case.dt)
            k += 1
        got = spl.cpu().numpy()
        diff = np.where(okb, np.abs(got - ref), 0.0)
        nwet = okb.sum(axis=(1, 2, 3))
        ms = (case.V1 * got).sum(axis=(1, 2, 3))
        out[str(mark)] = {
            "max_rel_err": (diff.max(axis=(1, 2, 3)) / span).tolist(),
            "rms_rel_err": (np.sqrt((diff * diff).sum(axis=(1, 2, 3)) / nwet) / span).tolist(),
            "content_rel_change": (np.abs(ms - m0) / np.abs(m0)).tolist(),
            "below_initial_min": (np.maximum(lo - np.where(okb, got, np.inf).min(axis=(1, 2, 3)), 0) / span).tolist(),
            "above_initial_max": (np.maximum(np.where(okb, got, -np.inf).max(axis=(1, 2, 3)) - hi, 0) / span).tolist(),
        }
        _log(f"after {mark} steps: max rel err {out[str(mark)]['max_rel_err']}")
    res = {
        "metric": "split vs the oracle: per-tracer max and RMS error relative to the tracer's initial range",
        "value": float(max(max(v["max_rel_err"]) for v in out.values())), "unit": "relative error",
        "n_gpus": 1, "steps": marks[-1], "warmup": 0, "higher_is_better": False,
        "dtype": "f32/f64 against the oracle's f64",
        "data": "synthetic (solid-body rotation, Gaussian blob and slotted cylinder)",
        "config": {"workload": "solid body 128x128x20, 2 tracers [Gaussian, slotted cylinder], Courant 0.4, closed "
                               "edges", "initial_span": span.tolist()},
        "checkpoints": out,
    }
    print(json.dumps(res), flush=True)
    return 0


def bench_diffusion(args) -> int:
    """Tracer diffusion of the whole stack on the regional grid: the along-layer Laplacian (K_h = 100 m^2/s, the
    anomaly field T - 10) then the implicit vertical solve (K_v = 1e-4 m^2/s, dz = max(h, 1e-6)), per step."""
    import torch

    from pyocean_amd import _native, ops, synth

    _log(f"generating regional {args.nx}x{args.ny}x{args.nz} nt={args.nt}")
    case = synth.regional(nx=args.nx, ny=args.ny, nz=args.nz, nt=args.nt, seed=0)
    nt, nz, ny, nx = case.C.shape
    act = case.ok.any(axis=0)
    h = np.where(act[None], case.V1 / case.V1.max() * 40.0, 0.0)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()       # noqa: E731
    cx, cy = d(np.ones((ny, nx + 1))), d(np.ones((ny + 1, nx)))
    okx, oky = d(np.ones((ny, nx + 1), np.uint8)), d(np.ones((ny + 1, nx), np.uint8))
    C, Tw, V, hd, actd = d(case.C), d(case.C - 10.0), d(case.V1), d(h), d(act.astype(np.uint8))
    dz = d(np.maximum(h, 1e-6).reshape(nz, -1))
    nu = d(np.full((nz + 1, ny * nx), 1e-4))
    out = torch.empty_like(C)
    ws = ops.Workspace()
    M = ny * nx

    def step():
        ops.diffuse_h(C, Tw, hd, actd, cx, cy, okx, oky, V, 100.0, case.dt, out=out)
        ops.vertical_implicit(out.view(nt, nz, M), dz, nu, None, case.dt, out=out.view(nt, nz, M), workspace=ws)

    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    _native.profile_enable(True)
    _native.profile_collect()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    _native.profile_enable(False)
    prof = _native.profile_collect()
    updates = nt * nz * ny * nx
    value = updates * args.steps / wall
    # compulsory bytes per tracer-cell: horizontal T, Tw read and out written (24 B) + h, V (16 B) / nt; vertical
    # u read and out written (16 B) + dz, nu (16 B) / nt
    b_h, b_v = 24.0 + 16.0 / nt, 16.0 + 16.0 / nt
    # the dominant kernel, the vertical solve: u read and out written + dz, cp, the lower coefficients and the
    # pivots (32 B) / nt per tracer-cell
    solve_ms = prof["vertical_solve"][0] / prof["vertical_solve"][1]
    solve_bytes = (16.0 + 32.0 / nt) * updates
    cpu = None
    if not args.no_cpu:
        from oracle import numpy_ref as R
        sy, sx, st = ny, nx, min(nt, 4)
        Cs = np.ascontiguousarray(case.C[:st, :, :sy, :sx])
        hs, acts, Vs = h[:, :sy, :sx], act[:sy, :sx], case.V1[:, :sy, :sx]
        Ms = sy * sx
        t1 = time.perf_counter()
        o = R.diffuse_h_ale(Cs, Cs - 10.0, hs, acts, np.ones((sy, sx + 1)), np.ones((sy + 1, sx)),
                            np.ones((sy, sx + 1), bool), np.ones((sy + 1, sx), bool), Vs, 100.0, case.dt)
        R.vertical_implicit(o.reshape(st, nz, Ms), np.maximum(hs, 1e-6).reshape(nz, Ms), np.full((nz + 1, Ms), 1e-4),
                            np.zeros(Ms), case.dt)
        tc = time.perf_counter() - t1
        cpu = {"value": st * nz * Ms / tc, "unit": "tracer-cell updates/s", "cores": 1, "kind": "port",
               "sample": f"{st} tracers of the same {sx}x{sy}x{nz} grid through the oracle's NumPy "
                         f"diffusion, along-layer then vertical ({tc:.1f} s)"}
    res = {
        "metric": "tracer-cell diffusion updates/s (along-layer Laplacian + implicit vertical, whole stack)",
        "value": value, "unit": "tracer-cell updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded regional grid)",
        "config": {"workload": f"regional {nx}x{ny}x{nz}, {nt} tracers, K_h 100 m2/s, K_v 1e-4 m2/s"},
        "roofline": {"kernel": "vertical_solve", "bound": "hbm", "achieved": solve_bytes / solve_ms / 1e6,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": solve_bytes / (solve_ms * 1e-3) / HBM_PEAK,
                     "traffic": None, "algorithmic_bytes": solve_bytes},
        "step_roofline": {"bytes_per_update": b_h + b_v, "achieved": value * (b_h + b_v) / 1e9,
                          "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": value * (b_h + b_v) / HBM_PEAK},
        "kernels_ms": {k: v[0] / v[1] for k, v in prof.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)
    return 0


# compulsory bytes per tracer-cell of the unlimited step's kernels (the shared Qx, Qy, V0, V1, Fz and ok, 41 B per
# cell, amortised over the tracers): RK3 stage 1 reads C, writes q1 and c1; stage 2 reads c1, C, q1, writes q2 and
# c2; stage 3 reads c2, C, q2, writes out; the 'pc' predictor reads C and writes c^{n+1/2}, the corrector reads it
# and C and writes out
RK_BYTES = {"rk3_stage1": 24.0, "rk3_stage2": 40.0, "rk3_stage3": 32.0, "pc_predictor": 16.0, "pc_corrector": 24.0}


def bench_rk3(args) -> int:
    """The unlimited step of TracerAdvection.advect on the regional grid in sigma form (V1 = V0, the continuity
    residual in Fz, so every stage runs the z sweep): SSP-RK3 or the 'pc' predictor-corrector, with the faces of
    --scheme."""
    import torch

    from pyocean_amd import _native, ops, synth

    _log(f"generating regional sigma {args.nx}x{args.ny}x{args.nz} nt={args.nt}")
    case = synth.regional(nx=args.nx, ny=args.ny, nz=args.nz, nt=args.nt, seed=0, sigma=True)
    nt, nz, ny, nx = case.C.shape
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()       # noqa: E731
    C, V0, V1, Qx, Qy, Fz = d(case.C), d(case.V0), d(case.V1), d(case.Qx), d(case.Qy), d(case.Fz)
    ok = d(case.ok.astype(np.uint8))
    out = torch.empty_like(C)
    ws = ops.Workspace()
    up = {"upwind3": 3, "upwind1": 1}.get(args.scheme, 0)
    flags = ops.flags_of(js=args.scheme == "weno5js", up=up)
    step_fn = ops.advect_pc if args.time == "pc" else ops.advect_rk3
    cur = [C, out]

    def step():
        step_fn(cur[0], V0, V1, Qx, Qy, Fz, ok, case.dt, flags=flags, vert=None if up else True, out=cur[1],
                workspace=ws)
        cur.reverse()

    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    _native.profile_enable(True)
    _native.profile_collect()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    _native.profile_enable(False)
    prof = _native.profile_collect()
    updates = nt * nz * ny * nx
    value = updates * args.steps / wall
    kms = {k: v[0] / v[1] for k, v in prof.items()}
    dom = max((k for k in kms if k in RK_BYTES), key=lambda k: kms[k])
    dom_bytes = (RK_BYTES[dom] + 41.0 / nt) * updates
    names = ("pc_predictor", "pc_corrector") if args.time == "pc" else ("rk3_stage1", "rk3_stage2", "rk3_stage3")
    b_step = sum(RK_BYTES[k] + 41.0 / nt for k in names)
    cpu = None
    if not args.no_cpu:
        from oracle import numpy_ref as R
        sy, st = min(ny, 64), 1
        Cs = np.ascontiguousarray(case.C[:st, :, :sy])
        sl = lambda a: np.ascontiguousarray(a[:, :sy])                    # noqa: E731
        Qys = np.ascontiguousarray(case.Qy[:, :sy + 1])
        Qys[:, sy] = 0.0
        run = R.advect_pc if args.time == "pc" else R.advect_rk3
        t1 = time.perf_counter()
        run(Cs, sl(case.V0), sl(case.V1), sl(case.Qx), Qys, sl(case.Fz), sl(case.ok), case.dt,
            z_variant=args.scheme != "weno5js", up_order=up)
        tc = time.perf_counter() - t1
        cpu = {"value": Cs.size / tc, "unit": "tracer-cell updates/s", "cores": 1, "kind": "port",
               "sample": f"{st} tracer of the first {sy} rows ({nx}x{sy}x{nz}) of the same grid through the "
                         f"oracle's NumPy {args.time} step ({tc:.1f} s)"}
    res = {
        "metric": f"tracer-cell updates/s of the unlimited {args.time} step ({args.scheme} faces, x, y and z sweeps)",
        "value": value, "unit": "tracer-cell updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded regional grid, sigma form)",
        "config": {"workload": f"regional sigma {nx}x{ny}x{nz}, {nt} tracers, {args.scheme}, {args.time}"},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": dom_bytes / kms[dom] / 1e6, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": dom_bytes / (kms[dom] * 1e-3) / HBM_PEAK, "traffic": None,
                     "algorithmic_bytes": dom_bytes},
        "step_roofline": {"bytes_per_update": b_step, "achieved": value * b_step / 1e9, "peak": HBM_PEAK / 1e9,
                          "unit": "GB/s", "frac": value * b_step / HBM_PEAK},
        "kernels_ms": kms,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)
    return 0


# algorithmic HBM bytes of the ALE layer-transport kernels per cell-layer (faces of one direction ~ cells): each
# donor cell is read by two faces, so a face reads one cell's worth of the cell fields
ALE_BYTES = {
    "face_thickness": 8.0 + 16.0,                     # h read, q_pos and q_neg written
    "transport_velocity": 3 * 8.0 + 8.0,              # u_new, u_old, w read, u^θ written
    "ppm_edges": 8.0 + 16.0,                          # h read, two edges written
    "ppm_init": 8.0 + 24.0 + 8.0,                     # u^θ, the donor's h and edges, the donor-side q
    "ppm_newton": 8.0 + 24.0 + 8.0,                   # the same per Newton pass
    "ppm_final": 8.0 + 24.0 + 16.0 + 8.0,             # u^θ, donor h and edges, both q, the flux written
    "outflow_over": 8.0 + 16.0 + 8.0,                 # h, Qx and Qy read (a face per cell each), r written
    "lagrangian": 8.0 + 16.0 + 24.0 + 1.0,            # h, Qx, Qy read; h†, V0, V1 and ok written
}


def bench_ale(args) -> int:
    """One ALE layer-transport step (the producer of the tracer step's Qx, Qy, V0, V1, ok) on the regional grid:
    transport thicknesses of both directions, transport velocities, PPM fluxes with the Newton correction, the
    outflow limiter (its check, and its passes where a layer would empty), the Lagrangian thickness and volumes."""
    import torch

    from pyocean_amd import _native, ale, synth

    _log(f"generating ALE inputs {args.nx}x{args.ny}x{args.nz}")
    a = synth.ale_inputs(args.nx, args.ny, args.nz, seed=0)
    d = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in vars(a).items() if isinstance(v, np.ndarray)}
    lt = ale.LayerTransport(Lu=d["Lu"], Lv=d["Lv"], area=d["area"], active=d["active"], dx_t=d["dx_t"],
                            dy_t=d["dy_t"], dt=a.dt, theta=a.theta)
    nz, ny, nx = a.h.shape
    state = {}

    def step(count=False):
        qe, qw = ale.face_thickness_axis(d["h"], d["H"], d["active"], True, 0)
        qn, qs = ale.face_thickness_axis(d["h"], d["H"], d["active"], True, 1)
        state["out"] = lt(d["h"], d["u_new"], d["u_old"], d["v_new"], d["v_old"], d["wet_u"], d["wet_v"], d["wu"],
                          d["wv"], d["Du"], d["Dv"], qe, qw, qn, qs, theta_bc=a.theta_bc, count_iterations=count)

    step(count=True)
    iters = lt.iterations
    passes = iters[0] + iters[1]
    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    if args.copy_calibration:                      # a device copy of known size for the traffic counters
        d["h"].clone()
        torch.cuda.synchronize()
    _native.profile_enable(True)
    _native.profile_collect()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    _native.profile_enable(False)
    prof = _native.profile_collect()
    cells = nz * ny * nx
    faces = nz * ny * (nx + 1)                     # per direction (the v-faces are as many to within 0.1 %)
    value = cells * args.steps / wall
    kern = {k: v[0] / args.steps for k, v in prof.items()}
    # launches that do work per step: every Newton pass launch after the last needed one returns at once
    work = {k: (passes if k == "ppm_newton" else prof[k][1] / args.steps) for k in prof}
    name = "ppm_newton"                            # the dominant kernel (its passes together)
    per_launch_ms = kern[name] / max(work[name], 1)
    b = ALE_BYTES[name] * faces
    achieved = b / (per_launch_ms * 1e-3) / 1e9
    step_bytes_ = sum(ALE_BYTES.get(k, 0.0) * cells * work[k] for k in kern)
    traffic = read_traffic(name, f"ale_regional{nx}x{ny}x{nz}")
    cpu = None
    if not args.no_cpu:
        from oracle import numpy_ref as R
        s = synth.ale_inputs(512, 512, nz, seed=0)
        t1 = time.perf_counter()
        R.ale_layer_transport(s.h, s.H, s.active, s.area, s.dx_t, s.dy_t, s.Lu, s.Lv, s.u_new, s.u_old, s.v_new,
                              s.v_old, s.wet_u, s.wet_v, s.wu, s.wv, s.Du, s.Dv, s.dt, s.theta, s.theta_bc)
        tc = time.perf_counter() - t1
        cpu = {"value": 512 * 512 * nz / tc, "unit": "cell-layer updates/s", "cores": 1, "kind": "port",
               "sample": f"one layer-transport step of the oracle (NumPy) on a 512x512x{nz} regional grid of the "
                         f"same construction ({tc:.1f} s)"}
    res = {
        "metric": "ALE layer-transport cell-layer updates/s (transport velocity, PPM flux + Newton, outflow "
                  "limiter, Lagrangian step; both face directions)",
        "value": value, "unit": "cell-layer updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded regional grid, z* layers, barotropic + baroclinic velocity)",
        "config": {"workload": f"ALE layer transport regional {nx}x{ny}x{nz}",
                   "newton_passes": {"u": iters[0], "v": iters[1]}, "n_limited": state["out"].n_limited},
        "roofline": {"kernel": name, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved * 1e9 / HBM_PEAK, "traffic": traffic, "per_launch_ms": per_launch_ms,
                     "algorithmic_bytes": b},
        "step_roofline": {"bytes_per_step": step_bytes_, "achieved": step_bytes_ / (wall / args.steps) / 1e9,
                          "unit": "GB/s", "frac": step_bytes_ / (wall / args.steps) / HBM_PEAK},
        "kernels_ms": kern,
        "working_launches_per_step": work,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)
    return 0


def bench_remap(args) -> int:
    """The vertical coordinate's regrid + remap step of the ALE cycle on the regional grid: the layers after one
    layer-transport step (h†) are regridded to z* (40 m nominal levels) and the tracer stack (--nt fields) is
    remapped onto them with PPM_H4."""
    import torch

    from pyocean_amd import _native, ale, synth, vertical

    _log(f"generating ALE inputs {args.nx}x{args.ny}x{args.nz}")
    a = synth.ale_inputs(args.nx, args.ny, args.nz, seed=0)
    d = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in vars(a).items() if isinstance(v, np.ndarray)}
    lt = ale.LayerTransport(Lu=d["Lu"], Lv=d["Lv"], area=d["area"], active=d["active"], dx_t=d["dx_t"],
                            dy_t=d["dy_t"], dt=a.dt, theta=a.theta)
    qe, qw = ale.face_thickness_axis(d["h"], d["H"], d["active"], True, 0)
    qn, qs = ale.face_thickness_axis(d["h"], d["H"], d["active"], True, 1)
    fl = lt(d["h"], d["u_new"], d["u_old"], d["v_new"], d["v_old"], d["wet_u"], d["wet_v"], d["wu"], d["wv"],
            d["Du"], d["Dv"], qe, qw, qn, qs, theta_bc=a.theta_bc)
    h_src = fl.h_dagger
    eta = h_src.sum(0) - d["H"]
    nz, ny, nx = a.h.shape
    rng = np.random.default_rng(2)
    C = torch.from_numpy(rng.normal(10.0, 2.0, (args.nt, nz, ny, nx))).cuda()
    dz = np.full(nz, 2000.0 / nz)
    ts = {}
    if args.rule == "hybrid":
        # a stratified column: T warm at the surface and cold below (e-folding 500 m), S = 35; the layer densities
        # are Wright's at 2e7 Pa, the targets span their range; the z* floor is a tenth of the nominal levels and
        # a sigma window covers 10 .. 40 m of water
        from pyocean_amd import eos
        zc = torch.cumsum(h_src.flip(0), 0).flip(0) - 0.5 * h_src
        ts = {"T": (4.0 + 16.0 * torch.exp(-zc / 500.0)).contiguous(), "S": torch.full_like(h_src, 35.0)}
        e = eos.WrightFullEOS()
        rho = e.potential_density(ts["T"], ts["S"], 2e7)
        params = {"dz_top": dz / 10.0, "rho_target": np.linspace(float(rho.min()), float(rho.max()), nz + 1),
                  "dsigma": np.full(nz, 1.0 / nz), "d_sigma": 10.0, "d_hybrid": 40.0, "keep_isopycnal": True}
        vc = vertical.VerticalCoordinate(H=d["H"], active=d["active"], rule="hybrid", params=params, eos=e)
    else:
        vc = vertical.VerticalCoordinate(H=d["H"], active=d["active"], rule="zstar", params={"dz_top": dz})
    state = {}

    def step():
        state["h"], state["C"] = vc.regrid_remap(eta, h_src, C, **ts)

    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    if args.copy_calibration:
        h_src.clone()
        torch.cuda.synchronize()
    _native.profile_enable(True)
    _native.profile_collect()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    _native.profile_enable(False)
    prof = _native.profile_collect()
    cells = nz * ny * nx
    value = args.nt * cells * args.steps / wall
    kern = {k: v[0] / args.steps for k, v in prof.items()}
    # remap: each field's values read and written once (16 B per field-cell), both thickness stacks once per
    # column shared by the fields (16 B per cell)
    b = (16.0 * args.nt + 16.0) * cells
    achieved = b / (kern["remap"] * 1e-3) / 1e9
    # the traffic and instruction counters on file were taken with the z* rule
    traffic = read_traffic("remap", f"remap_regional{nx}x{ny}x{nz}nt{args.nt}") if args.rule == "zstar" else None
    valu_busy = None
    cpath = ROOT / "profiles" / f"counters_remap_regional{nx}x{ny}x{nz}nt{args.nt}.txt"
    if cpath.exists() and args.rule == "zstar":
        ctr = dict(l.split()[:2] for l in cpath.read_text().splitlines() if l and not l.startswith("#"))
        # VALU wavefront instructions x 4 cycles over the SIMD-cycles of the measured launch time
        valu_busy = float(ctr["SQ_INSTS_VALU"]) * 4 / (1024 * 2.4e9 * kern["remap"] * 1e-3)
    cpu = None
    if not args.no_cpu:
        from oracle import cref
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
        os.environ["OMP_NUM_THREADS"] = str(threads)
        hs, hn = h_src.cpu().numpy(), state["h"].cpu().numpy()
        etan = eta.cpu().numpy()
        sample = np.ascontiguousarray(C.cpu().numpy())
        reps = 6
        t1 = time.perf_counter()
        for _ in range(reps):
            if args.rule == "hybrid":
                from oracle import numpy_ref as R
                rho_np = R.wright_density(ts["T"].cpu().numpy(), ts["S"].cpu().numpy(), 2e7)
                _, wet = cref.regrid_hybrid(a.H, etan, a.active, hs, rho_np, vc.params)
            else:
                _, wet = cref.regrid(a.H, etan, a.active, hs, "zstar", dz)
            cref.remap_columns(hs, sample, hn, cref.PPM_H4, wet)
        tc = time.perf_counter() - t1
        cpu = {"value": reps * args.nt * cells / tc, "unit": "field-cell remaps/s", "cores": threads, "kind": "port",
               "sample": f"the same {args.rule} regrid and {args.nt}-field remap, {reps} repetitions, PPM_H4, the "
                         f"oracle's C loops over columns (OpenMP) ({tc:.1f} s)"}
    res = {
        "metric": f"field-cell remaps/s ({'z*' if args.rule == 'zstar' else 'hybrid'} regrid + PPM_H4 conservative "
                  f"remap of the tracer stack)",
        "value": value, "unit": "field-cell remaps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded regional grid after one layer-transport step)",
        "config": {"workload": f"regrid_remap regional {nx}x{ny}x{nz}, {args.nt} fields, {args.rule}, PPM_H4"},
        "roofline": {"kernel": "remap", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved * 1e9 / HBM_PEAK, "traffic": traffic, "algorithmic_bytes": b,
                     "valu_busy_from_counters": valu_busy},
        "kernels_ms": kern,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)
    return 0


def cycle_cpu_baseline(nx, ny, nz, nt):
    """The same ALE tracer cycle through the oracle on one host core, on an nx x ny x nz grid of the same
    construction: the NumPy layer transport, the C loops of the FCT advection, the NumPy diffusion, and the C z*
    regrid and PPM_H4 remap."""
    from oracle import cref
    from oracle import numpy_ref as R
    from pyocean_amd import synth

    a = synth.ale_inputs(nx, ny, nz, seed=0)
    rng = np.random.default_rng(3)
    C0 = 10.0 + rng.normal(0.0, 1.0, (nt, nz, ny, nx)) * a.active
    act = a.active.astype(bool)
    cx = a.Lu / np.concatenate([a.dx_t[:, :1], 0.5 * (a.dx_t[:, 1:] + a.dx_t[:, :-1]), a.dx_t[:, -1:]], 1)
    cy = a.Lv / np.concatenate([a.dy_t[:1], 0.5 * (a.dy_t[1:] + a.dy_t[:-1]), a.dy_t[-1:]], 0)
    okx = np.zeros((ny, nx + 1), bool)
    okx[:, 1:-1] = act[:, 1:] & act[:, :-1]
    oky = np.zeros((ny + 1, nx), bool)
    oky[1:-1] = act[1:] & act[:-1]
    M = ny * nx
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    t0 = time.perf_counter()
    fl = R.ale_layer_transport(a.h, a.H, a.active, a.area, a.dx_t, a.dy_t, a.Lu, a.Lv, a.u_new, a.u_old, a.v_new,
                               a.v_old, a.wet_u, a.wet_v, a.wu, a.wv, a.Du, a.Dv, a.dt, a.theta, a.theta_bc)
    C1 = cref.advect_fct(C0, fl["V0"], fl["V1"], fl["Qx"], fl["Qy"], fl["ok"], a.dt)
    hd = fl["h_dagger"]
    C2 = R.diffuse_h_ale(C1, C1, hd, act, cx, cy, okx, oky, fl["V1"], 100.0, a.dt)
    C2 = R.vertical_implicit(C2.reshape(nt, nz, M), np.maximum(hd, 1e-6).reshape(nz, M), np.full((nz + 1, M), 1e-4),
                             np.zeros(M), a.dt).reshape(nt, nz, ny, nx)
    eta = hd.sum(0) - a.H
    h_new, wet = cref.regrid(a.H, eta, a.active, hd, "zstar", np.full(nz, 2000.0 / nz))
    cref.remap_columns(hd, C2, h_new, cref.PPM_H4, wet)
    tc = time.perf_counter() - t0
    return {"value": nt * nz * ny * nx / tc, "unit": "tracer-cell updates/s", "cores": threads, "kind": "port",
            "sample": f"one cycle of the oracle on a {nx}x{ny}x{nz} grid of the same construction with {nt} tracers "
                      f"({tc:.1f} s): NumPy layer transport and diffusion on one core, C advection, regrid and "
                      f"remap with OpenMP on {threads} threads"}


def bench_cycle(args) -> int:
    """One ALE tracer cycle on the device, every phase on its own kernels and nothing leaving HBM: the layer
    transport (transport thicknesses, PPM fluxes with the Newton correction, outflow limiter, Lagrangian step), the
    FCT advection of the --nt tracer stack, its along-layer and implicit vertical diffusion, and the z* regrid with
    the PPM_H4 remap of the stack."""
    import torch

    from pyocean_amd import _native, ale, ops, synth, vertical

    _log(f"generating ALE inputs {args.nx}x{args.ny}x{args.nz}")
    a = synth.ale_inputs(args.nx, args.ny, args.nz, seed=0)
    d = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in vars(a).items() if isinstance(v, np.ndarray)}
    nz, ny, nx = a.h.shape
    nt = args.nt
    rng = np.random.default_rng(3)
    C0 = torch.from_numpy(10.0 + rng.normal(0.0, 1.0, (nt, nz, ny, nx)) * a.active).cuda()
    lt = ale.LayerTransport(Lu=d["Lu"], Lv=d["Lv"], area=d["area"], active=d["active"], dx_t=d["dx_t"],
                            dy_t=d["dy_t"], dt=a.dt, theta=a.theta)
    vc = vertical.VerticalCoordinate(H=d["H"], active=d["active"], rule="zstar",
                                     params={"dz_top": np.full(nz, 2000.0 / nz)})
    act8 = d["active"].to(torch.uint8)
    cx = (d["Lu"] / torch.cat([d["dx_t"][:, :1], 0.5 * (d["dx_t"][:, 1:] + d["dx_t"][:, :-1]), d["dx_t"][:, -1:]],
                              1)).contiguous()
    cy = (d["Lv"] / torch.cat([d["dy_t"][:1], 0.5 * (d["dy_t"][1:] + d["dy_t"][:-1]), d["dy_t"][-1:]], 0)).contiguous()
    okx = torch.zeros((ny, nx + 1), dtype=torch.uint8, device="cuda")
    okx[:, 1:-1] = act8[:, 1:] & act8[:, :-1]
    oky = torch.zeros((ny + 1, nx), dtype=torch.uint8, device="cuda")
    oky[1:-1] = act8[1:] & act8[:-1]
    nu = torch.full((nz + 1, ny * nx), 1e-4, dtype=torch.float64, device="cuda")
    ws_fct, ws_v = ops.Workspace(), ops.Workspace()
    C1 = torch.empty_like(C0)
    C2 = torch.empty_like(C0)
    events = {}

    def mark(name):
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        events.setdefault(name, []).append(e)

    def cycle(timed=False):
        if timed:
            mark("start")
        qe, qw = ale.face_thickness_axis(d["h"], d["H"], d["active"], True, 0)
        qn, qs = ale.face_thickness_axis(d["h"], d["H"], d["active"], True, 1)
        fl = lt(d["h"], d["u_new"], d["u_old"], d["v_new"], d["v_old"], d["wet_u"], d["wet_v"], d["wu"], d["wv"],
                d["Du"], d["Dv"], qe, qw, qn, qs, theta_bc=a.theta_bc)
        if timed:
            mark("layer_transport")
        ops.advect_fct(C0, fl.V0, fl.V1, fl.Qx, fl.Qy, fl.ok, a.dt, out=C1, workspace=ws_fct)
        if timed:
            mark("advection")
        ops.diffuse_h(C1, C1, fl.h_dagger, act8, cx, cy, okx, oky, fl.V1, 100.0, a.dt, out=C2)
        dz = torch.clamp(fl.h_dagger, min=1e-6).reshape(nz, -1)
        ops.vertical_implicit(C2.view(nt, nz, -1), dz, nu, None, a.dt, out=C2.view(nt, nz, -1), workspace=ws_v)
        if timed:
            mark("diffusion")
        eta = fl.h_dagger.sum(0) - d["H"]
        h_new, Cn = vc.regrid_remap(eta, fl.h_dagger, C2)
        if timed:
            mark("regrid_remap")
        return h_new, Cn

    for _ in range(max(args.warmup, 1)):
        cycle()
    torch.cuda.synchronize()
    events.clear()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        h_new, Cn = cycle(timed=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    names = ["layer_transport", "advection", "diffusion", "regrid_remap"]
    prev = events["start"]
    phase = {}
    for nm in names:
        phase[nm] = float(np.mean([p.elapsed_time(q) for p, q in zip(prev, events[nm])]))
        prev = events[nm]
    cells = nz * ny * nx
    # content of the stack before and after (the cycle conserves Σ V c up to the open-edge-free round-off)
    V_new = h_new * d["area"][None]
    c0 = ops.content(C0, torch.where(d["active"][None], d["h"], torch.zeros_like(d["h"])) * d["area"][None])
    c1 = ops.content(Cn.contiguous(), V_new)
    rel = float(((c1 - c0).abs() / c0.abs()).max())
    cpu = None
    if not args.no_cpu:
        cpu = cycle_cpu_baseline(min(nx, 384), min(ny, 384), nz, nt)
    res = {
        "metric": "ALE tracer cycle tracer-cell updates/s (layer transport, FCT advection, diffusion, z* regrid + "
                  "PPM_H4 remap)",
        "value": nt * cells * args.steps / wall, "unit": "tracer-cell updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded regional grid, z* layers, barotropic + baroclinic velocity, 10 tracers)",
        "config": {"workload": f"ALE tracer cycle regional {nx}x{ny}x{nz}, {nt} tracers"},
        "phase_ms": phase,
        "content_rel_change": rel,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=1024)
    ap.add_argument("--nz", type=int, default=50)
    ap.add_argument("--nt", type=int, default=10)
    ap.add_argument("--precision", choices=("f64", "split"), default="f64",
                    help="f64: the plain FP64 drop-in (bit-identical to the reference); split: FP32 reconstruction "
                         "with FP64 accumulation and limiter")
    ap.add_argument("--rch", type=int, default=0, help="rows per chunk of the marching kernels (0: default)")
    ap.add_argument("--wpb", type=int, default=0, help="tracers (wavefronts) per workgroup (0: default)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-tracers", type=int, default=2)
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--rule", choices=("zstar", "hybrid"), default="zstar",
                    help="--form remap: the regrid rule (hybrid: Wright densities of a stratified column)")
    ap.add_argument("--scheme", choices=("weno5", "weno5js", "upwind3", "upwind1"), default="weno5",
                    help="--form rk3: the face scheme")
    ap.add_argument("--time", choices=("ssprk3", "pc"), default="ssprk3", help="--form rk3: the time stepping")
    ap.add_argument("--form", choices=("step", "batch", "serve", "precision", "diffusion", "ale", "remap", "cycle",
                                       "rk3"),
                    default="step",
                    help="step: one domain (the default workload); batch: 64 regional domains of uneven size in one "
                         "persistent launch; serve: the Poisson job stream through the continuous-batching loop; "
                         "precision: split against FP64 after --checkpoints steps; diffusion: the along-layer and "
                         "vertical tracer diffusion of the stack; ale: the ALE layer-transport step that "
                         "produces the advection's inputs; remap: the vertical regrid + remap of the stack; "
                         "cycle: the whole ALE tracer cycle (transport, advection, diffusion, regrid + remap); "
                         "rk3: the unlimited step with the z sweep (--scheme, --time)")
    ap.add_argument("--checkpoints", default="1,10,200,2000", help="precision form: step counts to compare at")
    ap.add_argument("--case", choices=("regional", "solid"), default="regional",
                    help="precision form: the regional case against the FP64 form, or the solid-body case against "
                         "the oracle")
    ap.add_argument("--mix", choices=("mixed", "split", "f64"), default="mixed",
                    help="batch form: member precisions (mixed: even members FP64, odd split)")
    ap.add_argument("--domains", type=int, default=64, help="batch/serve forms: number of regional domains")
    ap.add_argument("--autotune", action="store_true", help="batch form: sweep the launch shapes first")
    ap.add_argument("--same-shape", action="store_true", help="batch form: every domain --nx x --ny x --nz")
    ap.add_argument("--eager", action="store_true", help="batch form: eager launches instead of the captured graph")
    ap.add_argument("--jobs", type=int, default=500, help="serve form: number of jobs")
    ap.add_argument("--rate", type=float, default=20.0, help="serve form: mean arrivals per second")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="serve form: HBM budget")
    ap.add_argument("--workers", type=int, default=16, help="batch/serve forms: host threads generating grids")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="copy the tracer stack once (device to device) before the steps: a kernel of known bytes "
                         "for scaling the HBM counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: one GPU only; sharding the tracers, the domains or the (y, x) plane over several GPUs is "
              "not launched here, so --gpus must be 1", file=sys.stderr)
        return 2

    import torch

    from pyocean_amd import _native, ops, synth

    if args.form == "batch":
        return bench_batch(args)
    if args.form == "serve":
        return bench_serve(args)
    if args.form == "precision":
        return bench_precision_solid(args) if args.case == "solid" else bench_precision(args)
    if args.form == "diffusion":
        return bench_diffusion(args)
    if args.form == "ale":
        return bench_ale(args)
    if args.form == "remap":
        return bench_remap(args)
    if args.form == "cycle":
        return bench_cycle(args)
    if args.form == "rk3":
        return bench_rk3(args)

    if args.rch or args.wpb:
        from pyocean_amd import batch as _batch
        _batch.set_tuning(args.rch, args.wpb)
    print(f"[bench] generating regional {args.nx}x{args.ny}x{args.nz} nt={args.nt}", file=sys.stderr, flush=True)
    case = synth.regional(nx=args.nx, ny=args.ny, nz=args.nz, nt=args.nt, seed=0)
    dev = torch.device("cuda", 0)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)     # noqa: E731
    V0, V1, Qx, Qy = d(case.V0), d(case.V1), d(case.Qx), d(case.Qy)
    ok = d(case.ok.astype(np.uint8))
    C = d(case.C)
    out = torch.empty_like(C)
    ws = ops.Workspace()
    nt, nz, ny, nx = C.shape
    updates = nt * nz * ny * nx

    def step(a, b):
        return ops.advect_fct(a, V0, V1, Qx, Qy, ok, case.dt, out=b, workspace=ws, precision=args.precision)

    if args.copy_calibration:
        out.copy_(C)
        torch.cuda.synchronize()
        print(f"[bench] calibration copy of {C.numel() * 8} bytes", file=sys.stderr, flush=True)
    print(f"[bench] warmup {args.warmup}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        step(C, out)
        C, out = out, C
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    _native.profile_enable(True)
    _native.profile_collect()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record()
        step(C, out)
        ev[k][1].record()
        C, out = out, C
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    _native.profile_enable(False)
    prof = _native.profile_collect()
    step_ms = np.array([a.elapsed_time(b) for a, b in ev])
    # conservation of the last step: sum V1 c^{n+1} against sum V0 c^n (closed edges)
    m0 = ops.content(out, V0).cpu().numpy()
    m1 = ops.content(C, V1).cpu().numpy()
    print(f"[bench] {args.steps} steps in {wall:.3f} s", file=sys.stderr, flush=True)

    value = updates * args.steps / wall
    dom = max(prof, key=lambda k: prof[k][0])
    dom_ms = prof[dom][0] / prof[dom][1]
    per_tc, shared = KERNEL_BYTES[args.precision].get(dom, (float("nan"), 0.0))
    alg = updates * (per_tc + shared / nt)
    achieved = alg / (dom_ms * 1e-3)
    workload = f"regional{args.nx}x{args.ny}x{args.nz}nt{args.nt}" + ("" if args.precision == "f64" else "split")
    traffic = read_traffic(dom, workload)
    B = step_bytes(nt, args.precision)
    # the step's kernels (hipEvents on their stream), per step: the roofline of the step against the survey's
    # algorithmic bytes per update, and the HBM bytes the counters measured for the same kernels
    step_kernels = [k for k in prof if k.startswith("fct_")]
    kern_ms = sum(prof[k][0] for k in step_kernels) / args.steps
    step_traffic = None
    if all(read_traffic(k, workload) is not None for k in step_kernels):
        step_traffic = sum(read_traffic(k, workload) for k in step_kernels)
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "tracer-cell updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "p50_ms": float(np.percentile(step_ms, 50)),
        "p99_ms": float(np.percentile(step_ms, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "f64" else "f32/f64",
        "data": "synthetic (seeded: curvilinear metrics, land mask, z* layers, k^-3 stream function, 10 tracers)",
        "config": {"workload": f"regional {nx}x{ny}x{nz}, {nt} tracers, "
                               + ("plain FP64" if args.precision == "f64" else "FP32/FP64-split")
                               + " WENO5-Z + Zalesak FCT + SSP-RK3",
                   "nx": nx, "ny": ny, "nz": nz, "nt": nt, "dt": case.dt, "land_fraction": case.info["land"],
                   "wet_cells": case.info["wet_cells"],
                   "form": "plain_f64" if args.precision == "f64" else "split_f32_f64"},
        "wet_updates_per_s": value * case.info["wet_cells"] / (nz * ny * nx),
        "roofline": {"bound": "hbm", "kernel": "the FCT step (" + " + ".join(step_kernels) + ")",
                     "achieved": updates * B / (kern_ms * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": updates * B / (kern_ms * 1e-3) / HBM_PEAK, "traffic": step_traffic,
                     "bytes_per_update": B, "updates_per_launch": updates, "avg_launch_ms": kern_ms,
                     "basis": "algorithmic bytes per tracer-cell update of the three-pass step (72 + 99/nt FP64, "
                              "52 + 99/nt split) over the step's kernel time; traffic = FETCH_SIZE + WRITE_SIZE of "
                              "those kernels per step, scaled by the copy calibration"},
        "wall_roofline": {"bytes_per_update": B, "achieved": value * B / 1e9, "peak": HBM_PEAK / 1e9,
                          "unit": "GB/s", "frac": value * B / HBM_PEAK},
        "dominant_kernel_design_bytes": {"kernel": dom, "achieved": achieved / 1e9, "unit": "GB/s",
                                         "frac": achieved / HBM_PEAK, "traffic": traffic,
                                         "design_bytes_per_launch": alg, "avg_launch_ms": dom_ms,
                                         "basis": "the bytes this design's kernel must touch (its stored face "
                                                  "stacks included) over its hipEvent-timed launch"},
        "kernels_ms": {k: v[0] / v[1] for k, v in prof.items()},
        "last_step_content_rel_err": float(np.max(np.abs(m1 - m0) / np.abs(m0))),
    }
    if not args.no_cpu:
        print(f"[bench] CPU baseline ({args.cpu_tracers} tracers x {args.cpu_reps})", file=sys.stderr, flush=True)
        res["cpu_baseline"] = cpu_baseline(case, args.cpu_tracers, args.cpu_reps)
    else:
        res["cpu_baseline"] = None
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
