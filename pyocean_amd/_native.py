"""
ctypes binding of ``libocx.so`` (the C ABI declared in ``include/ocx.h``).

There is no fallback: if the library is missing or fails to load, ``lib()`` raises, so nothing in the package can
silently run anything but the HIP kernels.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import os

LIB_PATH = Path(os.environ.get("OCX_LIB", Path(__file__).resolve().parent / "lib" / "libocx.so"))

OPEN_W, OPEN_E, OPEN_S, OPEN_N = 1, 2, 4, 8
WENO_JS = 16
UNLIMITED = 32
UP3 = 64
UP1 = 128
CENTRED = 256

OCX_E_ARG, OCX_E_HIP, OCX_E_WORKSPACE, OCX_E_RANGE = -1, -2, -3, -4
OUTFLOW_NOOP, OUTFLOW_OK, OUTFLOW_ACCEPTED, OUTFLOW_EXCEEDS, OUTFLOW_EMPTY_DONOR, OUTFLOW_NEGATIVE = range(6)

# (name, restype, argtypes) of every exported symbol; the CPU tests check that each one is exported
P, I, U, D, Z, L = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint, ctypes.c_double, ctypes.c_size_t, ctypes.c_longlong
SIGNATURES = {
    "ocx_last_error": (ctypes.c_char_p, []),
    "ocx_version": (I, []),
    "ocx_create": (I, [I, P]),
    "ocx_destroy": (I, [P]),
    "ocx_ctx_stream": (P, [P]),
    "ocx_alloc": (I, [P, Z, P]),
    "ocx_free": (I, [P, P]),
    "ocx_h2d": (I, [P, P, P, Z]),
    "ocx_d2h": (I, [P, P, P, Z]),
    "ocx_sync": (I, [P]),
    "ocx_fct_workspace_bytes": (Z, [I, I, I, I]),
    "ocx_fct_split_workspace_bytes": (Z, [I, I, I, I]),
    "ocx_rk3_workspace_bytes": (Z, [I, I, I, I]),
    "ocx_advect_fct_f64": (I, [P, P, P, P, P, P, P, I, I, I, I, D, U, I, P, P, P, P, Z, P]),
    "ocx_advect_fct_split": (I, [P, P, P, P, P, P, P, I, I, I, I, D, U, I, P, P, P, P, Z, P]),
    "ocx_advect_fct_f64_noalloc": (I, [P, P, P, P, P, P, P, I, I, I, I, D, U, I, P, P, P, P]),
    "ocx_advect_rk3_f64": (I, [P, P, P, P, P, P, P, P, I, I, I, I, D, U, I, I, P, P, P, Z, P]),
    "ocx_advect_pc_f64": (I, [P, P, P, P, P, P, P, P, I, I, I, I, D, U, I, I, P, P, P, Z, P]),
    "ocx_faces_h_f64": (I, [P, P, P, P, P, I, I, I, I, U, P, P, P]),
    "ocx_faces_f64": (I, [P, P, P, L, I, I, L, U, P, P]),
    "ocx_open_faces_f64": (I, [P, P, P, P, I, I, I, I, I, U, P]),
    "ocx_upwind_faces_h_f64": (I, [P, P, P, P, I, I, I, I, U, P, P, P]),
    "ocx_rhs_f64": (I, [P, P, P, P, P, P, I, I, I, I, U, I, P, P]),
    "ocx_fct_coefficients_f64": (I, [P, P, P, P, P, P, D, I, I, I, I, P, P, P, Z, P]),
    "ocx_periodic_pad": (I, [P, P, Z, I, I, I, I, I, I, I, P]),
    "ocx_mix_num_f64": (I, [P, P, P, I, I, I, D, I, P, P]),
    "ocx_diffuse_h_f64": (I, [P, P, P, P, P, P, P, P, P, I, I, I, I, D, D, P, P]),
    "ocx_vertical_implicit_workspace_bytes": (Z, [I, L]),
    "ocx_vertical_implicit_f64": (I, [P, P, P, P, I, P, I, I, L, D, P, P, Z, P]),
    "ocx_selftest_div_f64": (I, [P, P, L, P, P]),
    "ocx_content_workspace_bytes": (Z, [I]),
    "ocx_content_f64": (I, [P, P, I, I, I, I, P, P, Z, P]),
    "ocx_transport_velocity_f64": (I, [P, P, P, P, P, P, I, L, D, D, P, P, P]),
    "ocx_face_transport_thickness_f64": (I, [P, P, P, I, I, I, I, I, P, P, P]),
    "ocx_ppm_edges_f64": (I, [P, P, I, I, I, I, D, P, P, P]),
    "ocx_ppm_transport_workspace_bytes": (Z, [I, I, I, I, I]),
    "ocx_ppm_transport_f64": (I, [P, P, P, P, P, P, P, P, I, I, I, I, D, I, P, P, Z, P]),
    "ocx_ppm_transport_iterations": (I, [P, I, I, I, I, I, P]),
    "ocx_limit_outflow_workspace_bytes": (Z, [I, I, I]),
    "ocx_limit_outflow_f64": (I, [P, P, P, P, P, I, I, I, D, D, I, I, I, I, P, Z, P, P, P, P, P]),
    "ocx_lagrangian_workspace_bytes": (Z, []),
    "ocx_lagrangian_thickness_f64": (I, [P, P, P, P, P, P, I, I, I, D, D, P, P, P, P, P, Z, P, P]),
    "ocx_remap_columns_f64": (I, [P, P, P, P, I, I, L, I, I, P, P]),
    "ocx_regrid_workspace_bytes": (Z, []),
    "ocx_regrid_f64": (I, [P, P, P, P, I, L, I, P, D, P, P, P, Z, P, P]),
    "ocx_regrid_hybrid_workspace_bytes": (Z, [I, L]),
    "ocx_regrid_hybrid_f64": (I, [P, P, P, P, P, I, L, P, P, P, P, D, D, D, I, P, P, P, Z, P, P]),
    "ocx_regrid_density_f64": (I, [I, P, P, P, D, P, L, P, P]),
    "ocx_upwind_transport_f64": (I, [P, P, P, P, P, I, I, I, I, P, P]),
    "ocx_open_face_transports_f64": (I, [P, P, P, P, P, P, P, P, P, P, P, P, P, I, I, I, U, D, P]),
    "ocx_freshwater_workspace_bytes": (Z, []),
    "ocx_freshwater_volumes_f64": (I, [P, P, D, P, D, P, I, I, I, P, P, P, Z, P, P]),
    "ocx_apply_freshwater_f64": (I, [P, I, I, I, I, I, I, I, P, P, P, D, P, P, Z, P, P]),
    "ocx_batch_create": (I, [P, I, I, I, I, P]),
    "ocx_batch_step": (I, [P, I, P]),
    "ocx_batch_graph_step": (I, [P, I, P]),
    "ocx_batch_destroy": (I, [P]),
    "ocx_set_tuning": (I, [I, I]),
    "ocx_profile_enable": (I, [I]),
    "ocx_profile_collect": (I, [I, P, P, P]),
}


class OcxDomain(ctypes.Structure):
    """ocx_domain of include/ocx.h (one member of a batch)."""
    _fields_ = [("C", P), ("C_out", P), ("V0", P), ("V1", P), ("Qx", P), ("Qy", P), ("inflow", P), ("ok", P),
                ("nt", I), ("nz", I), ("ny", I), ("nx", I), ("dt", D), ("flags", U), ("precision", I),
                ("workspace", P), ("workspace_bytes", Z)]


class OcxError(RuntimeError):
    """A libocx call returned an error code."""


_lib = None


def lib():
    """The loaded library with typed entry points (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"libocx.so is not built ({LIB_PATH}); run __graft_entry__.build() or "
                               "python -m pyocean_amd.build")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def profile_enable(on: bool) -> None:
    lib().ocx_profile_enable(int(bool(on)))


def profile_collect(max_kernels: int = 64) -> dict:
    """{kernel label: (total ms, launches)} of the launches recorded since the last collect."""
    names = (ctypes.c_char_p * max_kernels)()
    ms = (ctypes.c_double * max_kernels)()
    cnt = (ctypes.c_longlong * max_kernels)()
    n = lib().ocx_profile_collect(max_kernels, names, ms, cnt)
    if n < 0:
        check(n, "ocx_profile_collect")
    return {names[m].decode(): (ms[m], cnt[m]) for m in range(min(n, max_kernels))}


def check(rc: int, what: str) -> None:
    """Raise OcxError (ValueError for argument errors) when a call failed."""
    if rc == 0:
        return
    msg = lib().ocx_last_error().decode(errors="replace")
    if rc == OCX_E_ARG:
        raise ValueError(f"{what}: {msg}")
    raise OcxError(f"{what} failed ({rc}): {msg}")
