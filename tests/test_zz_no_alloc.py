"""
The allocation-free FCT entry (``ocx_advect_fct_f64_noalloc``) against ``ocx_advect_fct_f64`` through the C ABI:
bit for bit on two calls in a row, the second with a smaller input (the library's scratch keeps the first call's
contents), with open edges, inflow, the face-column output and the open-edge budget; a shape whose workspace exceeds
the scratch is rejected with OCX_E_WORKSPACE before any launch.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from pyocean_amd import _native
from tests import cases

pytestmark = pytest.mark.gpu


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _run(noalloc, case, inflow, fi, ws=None):
    L = _native.lib()
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()        # noqa: E731
    C, V0, V1, Qx, Qy = (d(case[k]) for k in ("C", "V0", "V1", "Qx", "Qy"))
    ok = d(case["ok"].astype(np.uint8))
    inf = None if inflow is None else d(inflow)
    nt, nz, ny, nx = C.shape
    out = torch.empty_like(C)
    fxi = torch.empty((nt, nz, ny), dtype=torch.float64, device=C.device)
    of = torch.empty((nt,), dtype=torch.float64, device=C.device)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    args = (_p(C), _p(V0), _p(V1), _p(Qx), _p(Qy), _p(ok), _p(inf), nt, nz, ny, nx, float(case["dt"]),
            int(case["open_mask"]), fi, _p(out), _p(fxi), _p(of))
    if noalloc:
        rc = L.ocx_advect_fct_f64_noalloc(*args, s)
    else:
        nb = L.ocx_fct_workspace_bytes(nt, nz, ny, nx)
        ws = torch.empty(nb, dtype=torch.uint8, device=C.device)
        rc = L.ocx_advect_fct_f64(*args, _p(ws), nb, s)
    _native.check(rc, "fct")
    torch.cuda.synchronize()
    return out.cpu().numpy(), fxi.cpu().numpy(), of.cpu().numpy()


def test_allocation_free_step_equals_the_workspace_step(gpu):
    big = cases.fused_fct_case(open_=True, nx=150, ny=120, nz=8)
    small = cases.fused_fct_case(open_=True, nx=40, ny=33, nz=3, seed=11)
    for name, case in (("150x120x8", big), ("40x33x3", small)):
        print(f"[gpu case] no-alloc {name}", flush=True)
        inflow = case["C"] * 0.5 + 1.0
        a = _run(True, case, inflow, 7)
        b = _run(False, case, inflow, 7)
        for x, y in zip(a, b):
            assert np.array_equal(x.view(np.int64), y.view(np.int64)), name


def test_allocation_free_step_rejects_shapes_beyond_its_scratch(gpu):
    L = _native.lib()
    nt, nz, ny, nx = 10, 50, 1024, 1024
    assert L.ocx_fct_workspace_bytes(nt, nz, ny, nx) > (64 << 20)
    v = ctypes.c_void_p(1)                                                   # never dereferenced: rejected first
    rc = L.ocx_advect_fct_f64_noalloc(v, v, v, v, v, v, None, nt, nz, ny, nx, 1.0, 0, -1, ctypes.c_void_p(2), None,
                                      None, None)
    assert rc == _native.OCX_E_WORKSPACE
    assert b"allocation-free" in L.ocx_last_error()
