"""The factored MRT transforms of the step kernels (csrc/mrt.hpp, host build of
the same header) against the moment basis of csrc/d3q19.hpp: on every unit
vector, mrt_moments must give column q of M and mrt_back row k of M (negated),
exactly, since every coefficient is a small integer.  Both maps are linear, so
this pins every coefficient of the factored forms."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r'''
#include "mrt.hpp"
#include <cstdio>
int main() {
  using namespace d3q19;
  int bad = 0;
  for (int q = 0; q < Q; ++q) {  // m = M e_q: column q
    float f[Q] = {}, m[Q];
    f[q] = 1.0f;
    nflow::mrt_moments(f, m);
    for (int k = 0; k < Q; ++k)
      if (m[k] != (float)mrow(k, q)) { std::printf("moments k=%d q=%d: %g vs %d\n", k, q, m[k], mrow(k, q)); ++bad; }
  }
  for (int k = 0; k < Q; ++k) {  // f -= M^T e_k: minus row k (conserved rows are not read)
    const bool conserved = k == 0 || k == 3 || k == 5 || k == 7;
    float f[Q] = {}, dl[Q] = {};
    dl[k] = 1.0f;
    nflow::mrt_back(f, dl);
    for (int q = 0; q < Q; ++q) {
      const float want = conserved ? 0.0f : -(float)mrow(k, q);
      if (f[q] != want) { std::printf("back k=%d q=%d: %g vs %g\n", k, q, f[q], want); ++bad; }
    }
  }
  std::printf("%d\n", bad);
  return bad != 0;
}
'''


def test_factored_transforms_equal_the_basis(tmp_path):
    src = tmp_path / "mrt.cpp"
    src.write_text(SRC)
    exe = tmp_path / "mrt"
    inc = os.path.join(ROOT, "nflow-red-poplar_amd", "csrc")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + inc,
                        str(src), "-o", str(exe)], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("host compiler cannot build against hip_runtime.h: " + r.stderr[-300:])
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
