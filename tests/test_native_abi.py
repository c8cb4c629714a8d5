"""The C-ABI library loads on the CPU, exports every symbol include/ocx.h declares, and rejects bad arguments on
the host before any launch (no GPU needed for any of this)."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

from pyocean_amd import _native

HEADER = Path(__file__).resolve().parents[1] / "include" / "ocx.h"


def _declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ocx_\w+)\s*\(", text)))


def test_header_and_binding_declare_the_same_symbols():
    assert _declared() == sorted(_native.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = _native.lib()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.ocx_version() == 1


def test_workspace_sizes_are_host_computations():
    lib = _native.lib()
    nt, nz, ny, nx = 10, 50, 1024, 1024
    b = lib.ocx_fct_workspace_bytes(nt, nz, ny, nx)
    S = nt * nz * ny * nx
    assert 6 * S * 8 <= b <= 6.1 * S * 8                    # 2 cell stacks + 2 x 2 face stacks + edge fluxes
    bs = lib.ocx_fct_split_workspace_bytes(nt, nz, ny, nx)
    assert 6 * S * 4 + S * 8 <= bs <= 6.1 * S * 4 + S * 8  # FP32 stages and fluxes, FP64 c_L
    assert lib.ocx_rk3_workspace_bytes(nt, nz, ny, nx) >= 3 * S * 8
    assert lib.ocx_fct_workspace_bytes(0, 1, 1, 1) == 0


def _fct(lib, **over):
    a = dict(C=1, V0=1, V1=1, Qx=1, Qy=1, ok=1, inflow=None, nt=1, nz=1, ny=8, nx=8, dt=1.0, flags=0, fi=-1,
             out=2, fxi=None, of=None, ws=3, wsb=1 << 40, stream=None)
    a.update(over)
    v = lambda x: None if x is None else ctypes.c_void_p(x)                 # noqa: E731  (never dereferenced)
    return lib.ocx_advect_fct_f64(v(a["C"]), v(a["V0"]), v(a["V1"]), v(a["Qx"]), v(a["Qy"]), v(a["ok"]),
                                  v(a["inflow"]), a["nt"], a["nz"], a["ny"], a["nx"], a["dt"], a["flags"], a["fi"],
                                  v(a["out"]), v(a["fxi"]), v(a["of"]), v(a["ws"]), a["wsb"], v(a["stream"]))


@pytest.mark.parametrize("over,msg", [
    (dict(nt=0), "must be positive"),
    (dict(C=None), "C is NULL"),
    (dict(ok=None), "ok is NULL"),
    (dict(out=1), "alias"),
    (dict(flags=1 << 8), "unknown flag"),
    (dict(fi=9), "flux_face_i"),
    (dict(fi=3), "fx_i is NULL"),
    (dict(ws=None), "workspace"),
    (dict(wsb=16), "workspace"),
])
def test_argument_errors_are_reported_before_any_launch(over, msg):
    lib = _native.lib()
    rc = _fct(lib, **over)
    assert rc in (_native.OCX_E_ARG, _native.OCX_E_WORKSPACE)
    assert msg in lib.ocx_last_error().decode()


def test_check_raises_value_error_for_argument_codes():
    lib = _native.lib()
    _fct(lib, nt=0)
    with pytest.raises(ValueError, match="positive"):
        _native.check(_native.OCX_E_ARG, "ocx_advect_fct_f64")


def test_content_needs_its_workspace():
    lib = _native.lib()
    assert lib.ocx_content_workspace_bytes(10) == 10 * 256 * 16
    assert lib.ocx_content_workspace_bytes(0) == 0
    v = ctypes.c_void_p
    rc = lib.ocx_content_f64(v(1), v(2), 10, 4, 8, 8, v(3), v(4), 16, None)
    assert rc == _native.OCX_E_WORKSPACE and "workspace" in lib.ocx_last_error().decode()
    rc = lib.ocx_content_f64(v(1), v(2), 10, 4, 8, 8, None, v(4), 1 << 20, None)
    assert rc == _native.OCX_E_ARG and "out is NULL" in lib.ocx_last_error().decode()


def test_planes_beyond_32_bit_offsets_are_rejected():
    lib = _native.lib()
    rc = _fct(lib, ny=20000, nx=20000)
    assert rc == _native.OCX_E_ARG and "too large" in lib.ocx_last_error().decode()


C_HOST = r"""
#include <stdio.h>
#include <string.h>
#include "ocx.h"

/* host-side calls only: the ABI version, workspace sizes, and an argument error with its message */
int main(void) {
    double one = 1.0, out = 0.0;
    unsigned char okv = 1;
    char ws[16];
    int rc;
    printf("version %d\n", ocx_version());
    printf("fct %zu split %zu rk3 %zu\n", ocx_fct_workspace_bytes(10, 50, 96, 80),
           ocx_fct_split_workspace_bytes(10, 50, 96, 80), ocx_rk3_workspace_bytes(10, 50, 96, 80));
    rc = ocx_advect_fct_f64(&one, &one, &one, &one, &one, &okv, NULL, 0, 1, 8, 8, 1.0, OCX_OPEN_W, -1, &out, NULL,
                            NULL, ws, sizeof ws, NULL);
    printf("rc %d %s\n", rc, rc == OCX_E_ARG && strstr(ocx_last_error(), "must be positive") ? "ok" : "bad");
    return 0;
}
"""


def test_a_c99_host_builds_against_the_header_and_links_the_library(tmp_path):
    """The boundary is plain C: a C99 program (gcc -std=c99 -pedantic, no HIP or C++ headers) includes
    include/ocx.h, links libocx.so and gets the same host-side answers as the Python binding."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    lib = _native.lib()
    src, exe = tmp_path / "host.c", tmp_path / "host"
    src.write_text(C_HOST)
    libdir = Path(_native.LIB_PATH).resolve().parent
    subprocess.run([gcc, "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", f"-I{HEADER.parent}", str(src),
                    f"-L{libdir}", f"-l:{Path(_native.LIB_PATH).name}", f"-Wl,-rpath,{libdir}", "-o", str(exe)],
                   check=True, capture_output=True)
    got = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    assert got[0] == f"version {lib.ocx_version()}"
    assert got[1] == (f"fct {lib.ocx_fct_workspace_bytes(10, 50, 96, 80)} "
                      f"split {lib.ocx_fct_split_workspace_bytes(10, 50, 96, 80)} "
                      f"rk3 {lib.ocx_rk3_workspace_bytes(10, 50, 96, 80)}")
    assert got[2] == f"rc {_native.OCX_E_ARG} ok"
