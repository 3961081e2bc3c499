"""Edge shapes on the device: lattices one node thick, a single node, the
smallest channel, extents that are not multiples of the 64/128-node storage
blocks, and work items that start inside the first x-z planes (the window of
16-bit storage clamps at node 0).  Every node of such a lattice is a boundary
node on at least one face."""
import numpy as np
import pytest

import nflow
from helpers import TOL_U_F16, cfg, fp32_parity, macro_errors, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1, 1), (1, 5, 3), (7, 1, 1), (3, 2, 1), (2, 9, 2), (65, 3, 2), (129, 2, 3), (5, 33, 7)]


@pytest.mark.parametrize("shape", SHAPES)
def test_cavity_edge_shapes_match_oracle(shape):
    nx, ny, nz = shape
    c = cfg(nx=nx, ny=ny, nz=nz, enable_les=1, smagorinsky=0.17, tau=0.6, init_lid_velocity=1)
    s = run_gpu(c, 25)
    o = run_oracle(c, 25)
    a, b = s.read_arF(), o.read_arF()
    # exact equality is not expected (FP32 vs FP64), but every population is close
    assert np.abs(a - b).max() < 2e-6, np.abs(a - b).max()
    t, d = s.node_type()
    to, do = o.node_type()
    assert np.array_equal(t, to) and np.array_equal(d, do)


@pytest.mark.parametrize("shape", [(3, 1, 1), (3, 4, 5), (3, 17, 2), (130, 3, 3)])
def test_channel_edge_shapes_match_oracle(shape):
    nx, ny, nz = shape
    c = cfg(nx=nx, ny=ny, nz=nz, casetype=nflow.CHANNEL, enable_les=1, smagorinsky=0.17, tau=0.7, inlet_peak=0.05,
            perturb=0.01, seed=3)
    s = run_gpu(c, 25)
    o = run_oracle(c, 25)
    assert np.abs(s.read_arF() - o.read_arF()).max() < 2e-6


@pytest.mark.parametrize("storage", [nflow.F16, nflow.BF16])
@pytest.mark.parametrize("shape", [(1, 1, 1), (65, 3, 2), (129, 2, 3), (40, 40, 40)])
def test_16bit_storage_edge_shapes(storage, shape):
    """16-bit storage (windowed buffer addressing) against FP32 storage of the
    same run: populations agree to the storage rounding of f - w."""
    nx, ny, nz = shape
    c = cfg(nx=nx, ny=ny, nz=nz, enable_les=1, smagorinsky=0.17, tau=0.6, init_lid_velocity=1)
    a = run_gpu(c, 30).read_arF()
    b = run_gpu(c, 30, storage=storage).read_arF()
    tol = 2e-4 if storage == nflow.F16 else 2e-3
    assert np.abs(a - b).max() < tol, np.abs(a - b).max()


def test_16bit_storage_with_tiles_and_every_block_size():
    # tiles split x, so work items start at arbitrary node ids; the result is
    # independent of tiles and launch shape
    c = cfg(nx=37, ny=11, nz=9, enable_les=1, smagorinsky=0.17, tau=0.6, init_lid_velocity=1)
    ref = run_gpu(c, 20, storage=nflow.F16).read_arF()
    inp = nflow.defaults(**c, storage=nflow.F16)
    slices = nflow.decomposition(inp, 3, 2, 2)
    s = run_gpu(c, 20, storage=nflow.F16, slices=slices)
    assert np.array_equal(s.read_arF(), ref)
    for tuning in [(64, 1, 0, 0), (128, 0, 2, 4), (512, 1, 8, 2)]:
        b = nflow.Batch([inp])
        b.set_tuning(*tuning)
        b.initialize()
        b.run(20)
        assert np.array_equal(b.read_arF(0), ref), tuning
        b.close()


def test_extents_beyond_the_window_range_are_rejected():
    with pytest.raises(nflow.NflowError):
        nflow.Solver(nflow.defaults(nx=8192, ny=1, nz=4096))
