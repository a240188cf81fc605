"""The C-ABI library loads and exports every function include/tinyad_amd.h declares
(no GPU needed: nothing is called that touches a device)."""
import ctypes
import os
import re

import tinyad_amd
from tinyad_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "tinyad_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tad_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_expected_set():
    assert declared_functions() == sorted(_lib.EXPORTED)


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_functions():
        assert hasattr(L, name), name


def test_version_and_error_string():
    assert "gfx950" in tinyad_amd.version()
    assert isinstance(tinyad_amd.lib().tad_last_error(), bytes)


def test_null_arguments_are_rejected_without_a_device():
    L = tinyad_amd.lib()
    assert L.tad_ctx_create(0, None) == 1
    assert L.tad_fn_destroy(None) == 0
    assert L.tad_fn_eval(None, None, 0, None) == 1


def test_struct_layouts_match_header(tmp_path):
    """Every struct the Python binding mirrors has, field by field, the size
    and offsets the C compiler gives it from include/tinyad_amd.h."""
    import subprocess
    mirrors = {"tad_solve_opts": _lib.SolveOpts, "tad_solve_stats": _lib.SolveStats,
               "tad_newton_opts": _lib.NewtonOpts, "tad_newton_stats": _lib.NewtonStats,
               "tad_device_ops": _lib.DeviceOps, "tad_device_residual_ops": _lib.DeviceResidualOps}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "tinyad_amd.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        a, b, v = line.split()
        got[(a, b)] = int(v)
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for field, _ in py._fields_:
            assert got[(cname, field)] == getattr(py, field).offset, (cname, field)
