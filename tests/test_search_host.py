"""Table searches of the kernels (csrc/clh_search.hpp) on the host: the
galloping walk and the bisection against a brute-force scan, and the
XCD-contiguous block order as a bijection (tests/search_check.cpp)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_search_helpers(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = tmp_path / "search_check"
    subprocess.run([cxx, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "cocoa-cosmolike-core_amd", "csrc"),
                    os.path.join(ROOT, "tests", "search_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    assert " bad 0" in out.stdout
