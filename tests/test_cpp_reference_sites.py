"""Reference-shaped C++ declarations compile unchanged against include/ibm.hpp
(tests/cpp/reference_sites.cpp: fd::dimension with fd::boundary objects and
unit literals, constexpr fd::domain / fd::mac / ib::spread / ib::interpolate,
unit values checked by static_assert against units.h's scales)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_reference_call_sites_compile():
    src = os.path.join(ROOT, "tests", "cpp", "reference_sites.cpp")
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-Werror", "-I",
                        os.path.join(ROOT, "include"), src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_api_check_program_compiles_as_plain_cpp():
    # the C++ API exercise program needs no HIP headers beyond the C ABI
    src = os.path.join(ROOT, "tests", "cpp", "api_check.cpp")
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), src],
                       capture_output=True, text=True)
    if "hip/hip_runtime" in r.stderr:
        pytest.skip("api_check.cpp uses the HIP runtime for its buffers")
    assert r.returncode == 0, r.stderr
