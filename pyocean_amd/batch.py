"""
Batched FCT step: many independent domains (ensemble members, regional domains of uneven size, FP64 or split
precision) advanced together by one persistent launch per stage (``ocx_batch_*``), optionally as a captured hipGraph.

Each domain keeps two state buffers in HBM and ping-pongs between them; ``Batch.state(d)`` is the current one.
A domain's result is bit-identical to what ``ops.advect_fct`` gives for it alone: the batch runs the same per-domain
kernel code, and nothing in a domain's arithmetic depends on the other members.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import _native, ops
from ._native import OcxDomain, check, lib


@dataclass
class Domain:
    """One batch member; all tensors on the GPU, shapes as in ``ops.advect_fct``."""
    C: torch.Tensor
    V0: torch.Tensor
    V1: torch.Tensor
    Qx: torch.Tensor
    Qy: torch.Tensor
    ok: torch.Tensor
    dt: float
    flags: int = 0
    precision: str = "f64"
    inflow: torch.Tensor | None = None
    C_alt: torch.Tensor | None = None        # the second state buffer (allocated by the batch when None)
    workspace: torch.Tensor | None = None    # uint8 device workspace (allocated by the batch when None)

    @property
    def cells(self) -> int:
        return self.C.numel()

    def workspace_bytes(self) -> int:
        nt, nz, ny, nx = self.C.shape
        L = lib()
        fn = L.ocx_fct_split_workspace_bytes if self.precision == "split" else L.ocx_fct_workspace_bytes
        return int(fn(nt, nz, ny, nx))

    def device_bytes(self) -> int:
        """HBM held by this member: its state, inputs and workspace."""
        t = [self.C, self.V0, self.V1, self.Qx, self.Qy, self.ok] + ([self.inflow] if self.inflow is not None else [])
        return 2 * self.C.numel() * 8 + sum(x.numel() * x.element_size() for x in t[1:]) + self.workspace_bytes()


class Batch:
    """The domains of one batch, their second state buffers and workspaces, and the device table."""

    def __init__(self, domains: list[Domain], rows_per_chunk: int = 0, waves_per_group: int = 0,
                 groups_per_cu: int = 0):
        if not domains:
            raise ValueError("a batch needs at least one domain")
        self.domains = list(domains)
        self.parity = 0
        self._bufs = []
        descs = (OcxDomain * len(domains))()
        for d, dom in enumerate(self.domains):
            nt, nz, ny, nx = dom.C.shape
            ops._need(dom.C, (nt, nz, ny, nx), "C")
            for t, s, nm in ((dom.V0, (nz, ny, nx), "V0"), (dom.V1, (nz, ny, nx), "V1"),
                             (dom.Qx, (nz, ny, nx + 1), "Qx"), (dom.Qy, (nz, ny + 1, nx), "Qy")):
                ops._need(t, s, nm)
            ok = ops._ok(dom.ok, (nz, ny, nx))
            if dom.inflow is not None:
                ops._need(dom.inflow, dom.C.shape, "inflow")
            if dom.precision not in ops.PRECISIONS:
                raise ValueError(f"domain {d}: precision {dom.precision!r}")
            if int(dom.flags) & ~31:
                raise ValueError(f"domain {d}: flags 0x{int(dom.flags):x}; a batch member takes only the open-edge "
                                 f"and WENO5-JS flags (0x1f)")
            other = dom.C_alt if dom.C_alt is not None else torch.empty_like(dom.C)
            ops._need(other, dom.C.shape, "C_alt")
            ws = dom.workspace
            if ws is None or ws.numel() < dom.workspace_bytes():
                ws = torch.empty(max(dom.workspace_bytes(), 256), dtype=torch.uint8, device=dom.C.device)
            self._bufs.append((other, ws, ok))
            p = lambda t: None if t is None else t.data_ptr()                # noqa: E731
            descs[d] = OcxDomain(p(dom.C), p(other), p(dom.V0), p(dom.V1), p(dom.Qx), p(dom.Qy), p(dom.inflow), p(ok),
                                 nt, nz, ny, nx, float(dom.dt), int(dom.flags),
                                 1 if dom.precision == "split" else 0, p(ws), ws.numel())
        self._descs = descs
        handle = ctypes.c_void_p()
        check(lib().ocx_batch_create(ctypes.cast(descs, ctypes.c_void_p), len(domains), int(rows_per_chunk),
                                     int(waves_per_group), int(groups_per_cu), ctypes.byref(handle)),
              "ocx_batch_create")
        self._h = handle

    def state(self, d: int) -> torch.Tensor:
        """The current state of domain d."""
        return self.domains[d].C if self.parity == 0 else self._bufs[d][0]

    def spare(self, d: int) -> torch.Tensor:
        """The other state buffer of domain d (the next step's output)."""
        return self._bufs[d][0] if self.parity == 0 else self.domains[d].C

    def workspace(self, d: int) -> torch.Tensor:
        """The workspace of domain d."""
        return self._bufs[d][1]

    def step(self, graph: bool = False, stream: torch.cuda.Stream | None = None) -> None:
        """One FCT step of every domain (eagerly, or as the captured graph of this parity)."""
        s = stream if stream is not None else torch.cuda.current_stream()
        fn = lib().ocx_batch_graph_step if graph else lib().ocx_batch_step
        check(fn(self._h, self.parity, ctypes.c_void_p(s.cuda_stream)),
              "ocx_batch_graph_step" if graph else "ocx_batch_step")
        self.parity ^= 1

    def updates(self) -> int:
        """Tracer-cell updates of one step (all cells, land included)."""
        return sum(d.cells for d in self.domains)

    def close(self) -> None:
        if self._h:
            lib().ocx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:                                                    # noqa: BLE001 (interpreter exit)
            pass


def set_tuning(rows_per_chunk: int = 0, waves_per_group: int = 0) -> None:
    """Launch shape of the single-domain entry points (0: default)."""
    check(_native.lib().ocx_set_tuning(int(rows_per_chunk), int(waves_per_group)), "ocx_set_tuning")
