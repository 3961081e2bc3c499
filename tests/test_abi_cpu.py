"""The C-ABI library without a GPU: it loads, exports every symbol that
include/nflow.h declares, and its host-only calls (parameter defaults,
decomposition, matrices, argument validation) behave."""
import os
import re
import subprocess

import numpy as np
import pytest

import nflow
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nflow.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nflow_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = nflow.lib()
    names = declared()
    assert len(names) >= 40
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(nflow.SIGNATURES), set(names) ^ set(nflow.SIGNATURES)


def test_input_defaults_are_the_reference_run():
    p = nflow.defaults()
    # main.cpp:59-85
    assert (p.h, p.nx, p.ny, p.nz, p.maxitr) == (10, 10, 10, 10, 199)
    assert p.enable_mrt == 1 and p.enable_les == 1 and p.smagorinsky == 0.0
    assert p.ref_den == 1.0 and tuple(p.ref_vel) == (0.1, 0.0, 0.0)
    assert abs(p.tau - (3 * 0.1 * 10 / 1000 + 0.5)) < 1e-15
    assert (p.w_eps, p.w_epsj, p.w_xx) == (3.0, -5.5, -0.5)
    assert p.casetype == nflow.CAVITY and p.init_lid_velocity == 1


def test_matrices():
    m = nflow.mrt_matrix()
    assert np.array_equal(m, oracle.mrt_matrix().astype(np.float32))
    inv = nflow.inverse_mrt(m)
    _, exact = oracle.mrt_inverse_exact(m.astype(np.float64))
    assert np.abs(inv - exact).max() < 1e-7
    with pytest.raises(nflow.NflowError) as e:
        nflow.inverse_mrt(np.ones((19, 19), dtype=np.float32))
    assert e.value.code == -1


def test_decomposition_errors():
    inp = nflow.defaults(nx=8, ny=8, nz=8)
    with pytest.raises(nflow.NflowError):
        nflow.decomposition(inp, 0, 1, 1)
    with pytest.raises(nflow.NflowError):
        nflow.decomposition(inp, 3, 3, 3, max_tiles=26)


def gpu_present():
    try:
        nflow.device_info()
        return True
    except nflow.NflowError:
        return False


def test_invalid_inputs_are_refused_before_the_device():
    for kw in (dict(tau=0.5), dict(nx=0), dict(casetype=7), dict(storage=9), dict(casetype=nflow.CHANNEL, nx=2),
               dict(ref_den=0.0), dict(nx=2048, ny=2048, nz=2048),
               dict(nx=1024, ny=2047, nz=1024)):  # 2^31 - 2^20 nodes: the first count refused
        with pytest.raises(nflow.NflowError) as e:
            nflow.Solver(nflow.defaults(**{**dict(nx=8, ny=8, nz=8), **kw}))
        assert e.value.code == -1, kw
    bad = np.array([[[0, 8], [0, 8], [0, 4]]], dtype=np.int32)
    with pytest.raises(nflow.NflowError) as e:
        nflow.Solver(nflow.defaults(nx=8, ny=8, nz=8), bad)
    assert e.value.code == -1
    overlap = np.array([[[0, 8], [0, 8], [0, 5]], [[0, 8], [0, 8], [3, 6]]], dtype=np.int32)
    with pytest.raises(nflow.NflowError):
        nflow.Solver(nflow.defaults(nx=8, ny=8, nz=8), overlap)


@pytest.mark.skipif(gpu_present(), reason="checks the no-device error path")
def test_no_device_is_an_error_code():
    with pytest.raises(nflow.NflowError) as e:
        nflow.Solver(nflow.defaults(nx=8, ny=8, nz=8))
    assert e.value.code == -3


def test_fast_division_is_exact(tmp_path):
    """make_fastdiv/fastdiv (lattice.hpp) against '/' for divisors up to 2^20
    and node ids across the 31-bit range (host build of the same header)."""
    src = tmp_path / "fd.cpp"
    src.write_text(r'''
#include "lattice.hpp"
#include <cstdio>
#include <random>
int main() {
  std::mt19937_64 g(1);
  long bad = 0, n = 0;
  for (uint32_t d = 1; d <= (1u << 20); d += (d < 4096 ? 1 : 97)) {
    nflow::FastDiv f = nflow::make_fastdiv(d);
    for (int i = 0; i < 200; ++i) {
      uint32_t x = (uint32_t)(g() & 0x7fffffffu);
      if (i < 4) x = (uint32_t[]){0u, d - 1, d, 0x7fffffffu}[i];
      bad += nflow::fastdiv(x, f) != x / d; ++n;
    }
  }
  std::printf("%ld %ld\n", bad, n);
  return bad != 0;
}
''')
    exe = tmp_path / "fd"
    inc = os.path.join(ROOT, "nflow-red-poplar_amd", "csrc")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + inc,
                        str(src), "-o", str(exe)], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("host compiler cannot build against hip_runtime.h: " + r.stderr[-300:])
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout


def header_defines():
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for name, val in re.findall(r"#define\s+(NFLOW_\w+)\s+(\(1 << \d+\)|\(-?\d+\)|-?\d+)", txt):
        m = re.fullmatch(r"\(1 << (\d+)\)", val)
        out[name] = 1 << int(m.group(1)) if m else int(val.strip("()"))
    return out


def test_case_numbers_and_node_flags_are_the_reference_values():
    """A reference caller copies Inputval.casetype and the node flags through
    unchanged: the header's, the binding's and the oracle's values are
    common.h's (tests/golden/common_h_constants.txt, extracted from it)."""
    ref = {}
    for line in open(os.path.join(ROOT, "tests", "golden", "common_h_constants.txt")):
        if line.strip() and not line.startswith("#"):
            k, v = line.split()
            ref[k] = int(v)
    h = header_defines()
    assert h["NFLOW_PIPE"] == ref["POISEUILLE"] == nflow.PIPE == oracle.PIPE
    assert h["NFLOW_CAVITY"] == ref["CAVITY"] == nflow.CAVITY == oracle.CAVITY
    assert h["NFLOW_CHANNEL"] == nflow.CHANNEL == oracle.CHANNEL
    assert len({h["NFLOW_PIPE"], h["NFLOW_CAVITY"], h["NFLOW_CHANNEL"]}) == 3
    for flag in ("FLUID", "FIXEDWALL", "MOVINGWALL", "INLET", "OUTLET", "STRUCT"):
        assert h["NFLOW_NODE_" + flag] == ref["NODE_" + flag] == getattr(nflow, "NODE_" + flag), flag
