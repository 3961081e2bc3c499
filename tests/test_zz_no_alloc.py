"""nflow_diagnostics_noalloc against nflow_diagnostics: the same sums bit for
bit on two calls in a row, the second on a smaller lattice (and the larger
one again after it, over the scratch the smaller call left), the same
divergence status, and a clear refusal above its 2^25-node limit."""
import numpy as np
import pytest

import nflow
from helpers import cfg, run_gpu

pytestmark = pytest.mark.gpu


def both(s):
    a = s.diagnostics()
    b = s.diagnostics_noalloc()
    return a, b


def same(a, b):
    return a[0] == b[0] and np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_noalloc_diagnostics_equal_the_allocating_ones():
    big = run_gpu(cfg(nx=96, ny=80, nz=72, enable_les=1, smagorinsky=0.17, tau=0.56, init_lid_velocity=1), 40)
    small = run_gpu(cfg(nx=40, ny=24, nz=32, casetype=nflow.CHANNEL, enable_les=1, smagorinsky=0.17, tau=0.56,
                        inlet_peak=0.05, perturb=0.01, seed=5), 30, storage=nflow.F16)
    a, b = both(big)
    assert same(a, b), (a, b)
    a2, b2 = both(small)  # second call, smaller input
    assert same(a2, b2), (a2, b2)
    assert b2[0] != b[0]
    big.run(7)  # and the larger lattice again, one step count later
    a3, b3 = both(big)
    assert same(a3, b3) and a3[2] != a[2]
    for s in (big, small):
        s.close()


def test_noalloc_diagnostics_report_divergence_and_refuse_oversized_lattices():
    # SRT at tau 0.5001 with a fast lid blows up within a few hundred steps
    bad = run_gpu(cfg(nx=24, ny=24, nz=24, enable_mrt=0, tau=0.5001, ref_vel=(0.3, 0.0, 0.0), init_lid_velocity=1), 0)
    for _ in range(40):
        bad.run(50)
        try:
            bad.diagnostics()
        except nflow.NflowError as e:
            assert e.code == nflow.EDIVERGED
            break
    else:
        pytest.fail("the near-inviscid SRT cavity did not diverge in 2000 steps")
    for f in (bad.diagnostics, bad.diagnostics_noalloc):
        with pytest.raises(nflow.NflowError) as e:
            f()
        assert e.value.code == nflow.EDIVERGED
    bad.close()
    # one plane more than 2^25 nodes (512 x 257 x 256), FP16: 2.6 GB
    huge = nflow.Solver(nflow.defaults(**cfg(nx=512, ny=257, nz=256), storage=nflow.F16))
    huge.initialize()
    with pytest.raises(nflow.NflowError, match="static scratch") as e:
        huge.diagnostics_noalloc()
    assert e.value.code == -1  # NFLOW_EINVAL
    huge.diagnostics()  # the allocating entry point has no such limit
    huge.close()
