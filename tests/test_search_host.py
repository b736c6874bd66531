"""Host checks of helpers that compile as plain C++: the table searches
(csrc/clh_search.hpp: galloping walk and bisection against a brute-force
scan, the XCD-contiguous block order as a bijection; tests/search_check.cpp),
the ln P run accumulator of the adjoint (csrc/clh_run.hpp against direct
sums, tiles in either order; tests/run_check.cpp) and the data-vector
placement (csrc/clh_place.hpp: block starts of every order, the Fourier
placement into a buffer of exactly the vector's size; tests/place_check.cpp,
built with AddressSanitizer) and the non-Limber FFTs (csrc/clh_fft.hpp: the
mixed-radix Stockham stages, c2r and r2c packing against direct sums;
tests/fft_check.cpp, AddressSanitizer).  The oracle, built with AddressSanitizer, runs
the Fourier-then-real sequence of the GPU Fourier test (one ell, block
order (0, 2, 1) among them)."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


ASAN = ["-fsanitize=address", "-fno-omit-frame-pointer", "-g"]


@pytest.mark.parametrize("name", ["search_check", "run_check", "place_check", "fft_check"])
def test_host_helpers(tmp_path, name):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = tmp_path / name
    extra = ASAN if name in ("place_check", "fft_check") else []
    subprocess.run([cxx, "-O2", "-std=c++17", *extra, "-I", os.path.join(ROOT, "cocoa-cosmolike-core_amd", "csrc"),
                    os.path.join(ROOT, "tests", name + ".cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    assert " bad 0" in out.stdout


ORACLE_SEQUENCE = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[2])
sys.path.insert(0, sys.argv[3])
import oracle as O
O._LIB = sys.argv[1]
from cosmolike_amd import synth
for nells, lmin, lmax, lmax_shear, ord in [(15, 30, 3000, 1000, (0, 1, 2)), (20, 20, 8000, 3000, (2, 1, 0)),
                                          (1, 100, 200, 10**6, (0, 2, 1))]:
    p = synth.full_3x2pt_problem()
    e = lmin * (lmax / lmin) ** (np.arange(nells + 1) / nells)
    o = O.Oracle().load(p)
    f = o.datavector_fourier(np.sqrt(e[:-1] * e[1:]), lmax_shear, ord)
    r = O.Oracle().load(p).datavector()
    assert f.size == nells * 110 and r.size == 4290 and np.isfinite(f).all() and np.isfinite(r).all()
print("sequence done")
"""


def test_oracle_fourier_sequence_under_asan(tmp_path):
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("no host C compiler")
    libasan = subprocess.run([cc, "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(libasan) or not os.path.exists(libasan):
        pytest.skip("no libasan")
    lib = tmp_path / "liboracle_asan.so"
    subprocess.run([cc, "-O1", "-fPIC", "-shared", "-std=gnu11", *ASAN, "-o", str(lib),
                    os.path.join(ROOT, "oracle", "cl_oracle.c"), "-lm"], check=True)
    script = tmp_path / "seq.py"
    script.write_text(ORACLE_SEQUENCE)
    env = dict(os.environ, LD_PRELOAD=libasan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               OMP_NUM_THREADS="4")
    out = subprocess.run([sys.executable, str(script), str(lib), os.path.join(ROOT, "oracle"),
                          os.path.join(ROOT, "cocoa-cosmolike-core_amd")],
                         capture_output=True, text=True, env=env, timeout=600)
    print(out.stdout[-2000:], out.stderr[-4000:])
    assert out.returncode == 0 and "sequence done" in out.stdout
    assert "AddressSanitizer" not in out.stderr
