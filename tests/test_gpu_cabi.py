"""
GPU tests through the C ABI alone - ctypes and NumPy, no torch - as a host without HIP of its own would drive the
library: a context, device buffers from ocx_alloc, copies, an FCT step and a remap on the context's stream, results
copied back and compared with the oracle bit for bit; argument errors come back as codes with a message.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle import cref
from pyocean_amd import _native, synth
from tests import cases

pytestmark = pytest.mark.gpu

P = ctypes.c_void_p


def _say(*a):
    print("[gpu case]", *a, flush=True)


class Ctx:
    def __init__(self):
        self.L = _native.lib()
        self.ctx = P()
        rc = self.L.ocx_create(0, ctypes.byref(self.ctx))
        assert rc == 0, self.L.ocx_last_error()
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        d = P()
        assert self.L.ocx_alloc(self.ctx, max(a.nbytes, 1), ctypes.byref(d)) == 0
        assert self.L.ocx_h2d(self.ctx, d, a.ctypes.data_as(P), a.nbytes) == 0
        self.bufs.append(d)
        return d

    def empty(self, nbytes):
        d = P()
        assert self.L.ocx_alloc(self.ctx, max(nbytes, 1), ctypes.byref(d)) == 0
        self.bufs.append(d)
        return d

    def get(self, d, shape, dtype=np.float64):
        out = np.empty(shape, dtype)
        assert self.L.ocx_d2h(self.ctx, out.ctypes.data_as(P), d, out.nbytes) == 0
        return out

    @property
    def stream(self):
        return P(self.L.ocx_ctx_stream(self.ctx))

    def close(self):
        for d in self.bufs:
            assert self.L.ocx_free(self.ctx, d) == 0
        assert self.L.ocx_destroy(self.ctx) == 0


def test_fct_step_through_the_c_abi_alone(gpu):
    _say("C ABI fct step")
    case = synth.regional(nx=96, ny=80, nz=8, nt=3, seed=4)
    nt, nz, ny, nx = case.C.shape
    c = Ctx()
    try:
        dC, dV0, dV1 = c.put(case.C), c.put(case.V0), c.put(case.V1)
        dQx, dQy, dok = c.put(case.Qx), c.put(case.Qy), c.put(case.ok.astype(np.uint8))
        dout = c.empty(case.C.nbytes)
        wsb = c.L.ocx_fct_workspace_bytes(nt, nz, ny, nx)
        dws = c.empty(wsb)
        rc = c.L.ocx_advect_fct_f64(dC, dV0, dV1, dQx, dQy, dok, None, nt, nz, ny, nx, ctypes.c_double(case.dt), 0, -1,
                                    dout, None, None, dws, ctypes.c_size_t(wsb), c.stream)
        assert rc == 0, c.L.ocx_last_error()
        assert c.L.ocx_sync(c.ctx) == 0
        got = c.get(dout, case.C.shape)
    finally:
        c.close()
    assert np.array_equal(got, cref.advect_fct(case.C, case.V0, case.V1, case.Qx, case.Qy, case.ok, case.dt))


def test_remap_through_the_c_abi_alone(gpu):
    _say("C ABI remap")
    hs, f, ht, wet = cases.remap_case(77, n0=30, n1=30)
    nf, n0 = f.shape[:2]
    M = hs[0].size
    c = Ctx()
    try:
        dh, df, dt_, dw = c.put(hs), c.put(f), c.put(ht), c.put(wet.astype(np.uint8))
        dout = c.empty(f.nbytes)
        rc = c.L.ocx_remap_columns_f64(dh, df, dt_, dw, n0, n0, M, nf, 2, dout, c.stream)
        assert rc == 0, c.L.ocx_last_error()
        got = c.get(dout, f.shape)
    finally:
        c.close()
    assert np.array_equal(got, cref.remap_columns(hs, f, ht, 2, wet), equal_nan=True)


def test_c_abi_argument_errors_return_codes(gpu):
    _say("C ABI errors")
    L = _native.lib()
    ctx = P()
    assert L.ocx_create(-1, ctypes.byref(ctx)) == _native.OCX_E_ARG
    assert b"device -1" in L.ocx_last_error()
    c = Ctx()
    try:
        assert c.L.ocx_h2d(c.ctx, None, None, 8) == _native.OCX_E_ARG
        d = c.empty(64)
        rc = c.L.ocx_remap_columns_f64(d, d, d, d, 4, 4, 1, 1, 7, d, c.stream)
        assert rc == _native.OCX_E_ARG and b"scheme = 7" in c.L.ocx_last_error()
    finally:
        c.close()
