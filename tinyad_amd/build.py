"""Builds libtinyad_amd.so in-tree with hipcc for gfx950.

Usage: python -m tinyad_amd.build [--force]
Each source is compiled to an object under _build/ in parallel, then linked.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "libtinyad_amd.so")
SOURCES = ["tad_element.hip", "tad_energy.hip", "tad_linalg.hip", "tad_dense.hip", "tad_basis.hip", "tad_adjoint.hip", "tad_api.cpp"]
HEADERS = ["tad_energies.hpp", "tad_elements.hpp", "tad_internal.hpp"]
# public headers the library's sources also compile (device code shared with
# element kernels built from user energies)
PUBLIC_HEADERS = ["tinyad_amd.h", "tinyad_amd/device/Dual.hpp", "tinyad_amd/device/Projection.hpp",
                  "tinyad_amd/device/Scatter.hpp", "tinyad_amd/device/Modes.hpp", "tinyad_amd/device/Kernels.hpp"]
ARCH = os.environ.get("TAD_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
         # packed FP32 instructions stay disabled in all device code
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def _compile(src: str, force: bool) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    srcp = os.path.join(CSRC, src)
    deps = ([srcp] + [os.path.join(CSRC, h) for h in HEADERS]
            + [os.path.join(ROOT, "include", h) for h in PUBLIC_HEADERS])
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= max(os.path.getmtime(p) for p in deps):
        return obj
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = ["hipcc"] + FLAGS + lang + ["-c", srcp, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-8000:]}")
    return obj


def build(force: bool = False, jobs: int = 4) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), SOURCES))
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < max(os.path.getmtime(o) for o in objs):
        tmp = LIB + ".tmp"
        cmd = ["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{r.stderr[-8000:]}")
        os.replace(tmp, LIB)
    return LIB


def build_device_energy_libs(force: bool = False) -> list:
    """Compiles the device-energy test plugins (tests/cpp/*.hip: energies as
    generic functors through include/tinyad_amd/DeviceElements.hh) with hipcc
    into tests/cpp/bin/libtad_test_<name>.so."""
    src_dir = os.path.join(ROOT, "tests", "cpp")
    out_dir = os.path.join(src_dir, "bin")
    os.makedirs(out_dir, exist_ok=True)
    deps = [os.path.join(ROOT, "include", h)
            for h in PUBLIC_HEADERS + ["tinyad_amd/DeviceElements.hh", "tinyad_amd/device/HyperDual.hpp"]]
    outs = []
    for src in sorted(f for f in os.listdir(src_dir) if f.endswith(".hip")):
        srcp = os.path.join(src_dir, src)
        lib = os.path.join(out_dir, "libtad_test_" + os.path.splitext(src)[0] + ".so")
        if force or not os.path.exists(lib) or os.path.getmtime(lib) < max(os.path.getmtime(p) for p in deps + [srcp]):
            cmd = ["hipcc"] + FLAGS + ["-shared", "-x", "hip", srcp, "-o", lib + ".tmp"]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-8000:]}")
            os.replace(lib + ".tmp", lib)
        outs.append(lib)
    return outs


def build_cpp_tests() -> list:
    """Compiles the C++ front-end test programs (tests/cpp/*.cc) with g++
    against include/, the library and the device-energy test plugin; binaries
    go to tests/cpp/bin/."""
    src_dir = os.path.join(ROOT, "tests", "cpp")
    out_dir = os.path.join(src_dir, "bin")
    os.makedirs(out_dir, exist_ok=True)
    outs = []
    for src in sorted(f for f in os.listdir(src_dir) if f.endswith(".cc")):
        exe = os.path.join(out_dir, os.path.splitext(src)[0])
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), os.path.join(src_dir, src),
               "-L" + HERE, "-ltinyad_amd", "-L" + out_dir, "-ltad_test_device_energies",
               "-Wl,-rpath,$ORIGIN/../../../tinyad_amd", "-Wl,-rpath,$ORIGIN", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed for {src}:\n{r.stderr[-4000:]}")
        outs.append(exe)
    return outs


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
    print(build_device_energy_libs(force="--force" in sys.argv))
    print(build_cpp_tests())
