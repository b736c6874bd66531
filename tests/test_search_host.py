"""Host checks of device helpers that compile as plain C++: the table
searches (csrc/clh_search.hpp: galloping walk and bisection against a
brute-force scan, the XCD-contiguous block order as a bijection;
tests/search_check.cpp) and the ln P run accumulator of the adjoint
(csrc/clh_run.hpp against direct sums, tiles in either order;
tests/run_check.cpp)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("name", ["search_check", "run_check"])
def test_host_helpers(tmp_path, name):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = tmp_path / name
    subprocess.run([cxx, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "cocoa-cosmolike-core_amd", "csrc"),
                    os.path.join(ROOT, "tests", name + ".cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    assert " bad 0" in out.stdout
