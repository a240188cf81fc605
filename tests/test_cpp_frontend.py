"""Runs the C++ front-end program (tests/cpp/test_frontend.cc, built by
__graft_entry__.build()) on the GPU: NewtonTest / ScalarFunctionTest /
ExceptionTest cases restated against include/tinyad_amd/ScalarFunction.hh."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "bin", "test_frontend")


@pytest.mark.gpu
def test_cpp_frontend_cases():
    assert os.path.exists(EXE), "build the C++ tests first (__graft_entry__.build())"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed" in r.stdout
