"""
pyocean_amd - pyOCEAN's 3-D tracer advection (WENO5-Z + Zalesak FCT + SSP-RK3 on the masked C-grid) as
hand-written HIP kernels for the MI355X (gfx950), behind the reference's ``TracerAdvection`` interface.

* ``pyocean_amd.tracers.TracerAdvection`` - the drop-in component (NumPy in, NumPy out);
* ``pyocean_amd.ops`` - the same operations on device-resident torch tensors;
* ``pyocean_amd.synth`` - seeded benchmark workloads;
* ``pyocean_amd.stack`` - the tracer stack's layout and its open-edge inflow values;
* ``include/ocx.h`` / ``pyocean_amd/lib/libocx.so`` - the C ABI and its library.
"""
from .periodic import NONE, TRACER_HALO, Periodic  # noqa: F401

__all__ = ["Periodic", "NONE", "TRACER_HALO"]
