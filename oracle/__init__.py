"""
The oracle: a CPU restatement of pyOCEAN's tracer-advection arithmetic (WENO5-Z/JS faces, Zalesak FCT, SSP-RK3).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py``
may import anything from this package, and only as the checker (or, for ``bench.py``, as the timed CPU baseline).
The product path (``pyocean_amd``) never imports it and fails loudly when its HIP library is missing.

Two independent forms:

* ``numpy_ref`` - vectorised NumPy, written from the reference's NumPy oracles
  (``core/dynamics/sigma_baroclinic.py:367-404, 581-629``; ``core/components/tracers.py:139-151, 178-193,
  256-340``);
* ``loops.c`` (``cref``) - plain C loops, written from the reference's compiled kernels
  (``core/dynamics/sigma_baroclinic.py:204-233, 265-364, 407-530, 856-875``), built with ``-ffp-contract=off``
  so that every product and sum is rounded on its own, as Numba (no fastmath) does.

Both keep the reference's operation order and IEEE divisions, so on the same inputs they agree bit for bit.
How each function is pinned (closed forms, the reference's 1-D known-answer table, invariants, the reference's
test thresholds) is stated in ``tests/test_oracle.py``.
"""
