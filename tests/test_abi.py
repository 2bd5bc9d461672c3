"""C-ABI checks that need no GPU: the library loads, exports every function
include/hitx.h declares, the ctypes signatures cover them, and the ctypes
descriptor has the C struct's layout (probed with gcc)."""

import ctypes
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "hitx.h"


def declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|size_t|const char\*)\s+(hitx_\w+)\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "hitx_rk3_stage" in names and "hitx_rhs" in names and len(names) >= 15


def test_library_exports_every_declared_symbol():
    from hit_solver_amd._lib import SIGNATURES, lib

    h = lib()
    for name in declared():
        assert hasattr(h, name), f"{name} not exported"
        assert name in SIGNATURES, f"{name} has no ctypes signature"
    assert h.hitx_abi_version() == 1
    nm = subprocess.run(["nm", "-D", "--defined-only", str(ROOT / "hit_solver_amd" / "libhitx.so")],
                        capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (hitx_\w+)", nm))
    assert set(declared()) <= exported


def test_plan_create_rejects_bad_descriptor_without_gpu():
    """Argument validation happens before any device call."""
    from hit_solver_amd._lib import HitxDesc, check, lib

    d = HitxDesc()
    d.abi_version = 1
    d.nx, d.ny = 4, 64  # too small
    h = ctypes.c_void_p()
    with pytest.raises(ValueError, match="nx and ny"):
        check(lib().hitx_plan_create(ctypes.byref(d), ctypes.byref(h)))
    d.nx = 64
    d.abi_version = 99
    with pytest.raises(ValueError, match="ABI"):
        check(lib().hitx_plan_create(ctypes.byref(d), ctypes.byref(h)))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_descriptor_layout_matches_c(tmp_path):
    from hit_solver_amd._lib import HitxDesc

    fields = [f[0] for f in HitxDesc._fields_]
    src = tmp_path / "probe.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "hitx.h"\nint main(void){\n'
        '  printf("%zu\\n", sizeof(hitx_desc));\n'
        + "".join(f'  printf("%zu\\n", offsetof(hitx_desc, {f}));\n' for f in fields)
        + "  return 0;\n}\n"
    )
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), "-o", str(exe), str(src)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert vals[0] == ctypes.sizeof(HitxDesc)
    for f, off in zip(fields, vals[1:]):
        assert getattr(HitxDesc, f).offset == off, f


def test_backend_selection_errors():
    from hit_solver_amd.backend import select_array_module

    with pytest.raises(ValueError):
        select_array_module("opencl")
    with pytest.raises(RuntimeError):
        select_array_module("cupy")
