"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the host-only code:
the library's decomposition API, moment basis, VTU/CSV writers and
checkpoint-header parser (nflow-red-poplar_amd/csrc/setup.cpp, io.cpp), and
the FP64 oracle (oracle/lbm_oracle.c).  Each is compiled with g++/gcc and the
sanitizers, then driven by tests/sanitize/*_driver.* over edge-case inputs;
any sanitizer report fails the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
           OMP_NUM_THREADS="2")


def _need(tool):
    if shutil.which(tool) is None:
        pytest.skip(f"{tool} not installed")


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    assert r.returncode == 0, f"{' '.join(cmd)}\n{r.stdout[-3000:]}\n{r.stderr[-3000:]}"
    return r


def test_host_library_code_under_sanitizers(tmp_path):
    _need("g++")
    pkg = os.path.join(ROOT, "nflow-red-poplar_amd")
    exe = str(tmp_path / "host_driver")
    _run(["g++", "-std=c++17", *SAN, "-I", os.path.join(ROOT, "include"), "-I", os.path.join(pkg, "csrc"),
          os.path.join(ROOT, "tests", "sanitize", "host_driver.cpp"), os.path.join(pkg, "csrc", "setup.cpp"),
          os.path.join(pkg, "csrc", "io.cpp"), "-o", exe])
    r = _run([exe, str(tmp_path)], env=ENV, timeout=300)
    assert "host ok" in r.stdout


def test_oracle_under_sanitizers(tmp_path):
    _need("gcc")
    exe = str(tmp_path / "oracle_driver")
    _run(["gcc", "-std=c11", *SAN, "-fopenmp", "-I", os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests", "sanitize", "oracle_driver.c"), os.path.join(ROOT, "oracle", "lbm_oracle.c"),
          "-lm", "-o", exe])
    r = _run([exe], env=ENV, timeout=600)
    assert "oracle ok" in r.stdout
