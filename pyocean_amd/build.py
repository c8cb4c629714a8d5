"""
Build recipe of the HIP library ``pyocean_amd/lib/libocx.so`` (gfx950, in-tree, no JIT cache).

Flags that matter:
* ``-ffp-contract=off``: no fused multiply-add, so the FP64 results equal the reference's arithmetic bit for bit;
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 instructions;
* ``--offload-arch=gfx950`` only.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "lib" / "libocx.so"
SOURCES = ["ocx_capi.hip", "ocx_ale.hip", "ocx_remap.hip", "ocx_ctx.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function"]


def build(verbose: bool = False, extra: list[str] | None = None) -> Path:
    """Compile every HIP source into libocx.so (always recompiles)."""
    LIB.parent.mkdir(exist_ok=True)
    tmp = LIB.with_suffix(f".so.tmp{os.getpid()}")
    cmd = [HIPCC, *FLAGS, *(extra or []), *[str(CSRC / s) for s in SOURCES], "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
