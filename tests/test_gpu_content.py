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
