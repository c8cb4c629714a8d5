"""
GPU tests of the content sums (sum_k,j,i V c per tracer): compensated and in a fixed order, so they agree with an
exactly rounded sum to about one ulp and repeat bit for bit.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from pyocean_amd import ops

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu case]", *a, flush=True)


@pytest.mark.parametrize("shape", [(3, 7, 33, 65), (2, 50, 257, 300), (1, 1, 1, 1)])
def test_content_is_near_exact_and_repeatable(gpu, shape):
    _say("content", shape)
    rng = np.random.default_rng(11)
    nt, nz, ny, nx = shape
    c = rng.uniform(-1.0, 1.0, shape) * 10.0 ** rng.integers(-3, 4, shape)
    V = rng.uniform(0.0, 1e6, (nz, ny, nx))
    cd, Vd = torch.from_numpy(c).cuda(), torch.from_numpy(V).cuda()
    got = ops.content(cd, Vd).cpu().numpy()
    again = ops.content(cd, Vd).cpu().numpy()
    assert np.array_equal(got, again)
    for t in range(nt):
        exact = math.fsum((V * c[t]).ravel())            # the products are rounded as the kernel rounds them
        scale = math.fsum(np.abs(V * c[t]).ravel())
        assert abs(got[t] - exact) <= 4e-16 * scale, (t, got[t], exact)


def test_numerical_mixing_matches_the_oracle_bit_for_bit(gpu):
    _say("mix_num")
    from oracle import numpy_ref as R
    from pyocean_amd import synth
    case = synth.regional(nx=90, ny=70, nz=6, nt=2, seed=5)
    c = np.stack([case.C[0], case.C[0] ** 2])
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()       # noqa: E731
    out = ops.advect_fct(d(c), d(case.V0), d(case.V1), d(case.Qx), d(case.Qy), d(case.ok.astype(np.uint8)), case.dt)
    got = ops.numerical_mixing(out[0], out[1], d(case.V1), case.dt)
    host = out.cpu().numpy()
    ref = R.numerical_mixing(host[0], host[1], case.V1, case.dt)
    assert np.array_equal(got.cpu().numpy(), ref)
    acc = got.clone()
    ops.numerical_mixing(out[0], out[1], d(case.V1), case.dt, out=acc, accumulate=True)
    assert np.array_equal(acc.cpu().numpy(), ref + ref)
