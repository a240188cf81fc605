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
