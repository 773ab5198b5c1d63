"""Builds tests/capi/capi_host_test.cpp with g++ against include/cogconv.h and the
built libcogconv.so, and runs it: the C ABI called from C++ directly, host-only."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT


def test_cpp_program_calls_the_c_abi(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    lib_dir = os.path.join(ROOT, "cct-core_amd")
    exe = str(tmp_path / "capi_host_test")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "capi", "capi_host_test.cpp"), "-L", lib_dir, "-lcogconv",
                    "-Wl,-rpath," + lib_dir, "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "checks passed" in r.stdout
