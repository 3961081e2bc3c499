"""Parity at the benchmark sizes: short runs against the FP64 oracle (with the
FP32-oracle round-off allowance) and size-independent properties over longer
runs (mass conservation, mirror symmetry, determinism)."""
import numpy as np
import pytest

import nflow
from helpers import cfg, fp32_parity, gpu_input, run_gpu

pytestmark = pytest.mark.gpu


def test_cavity_128_mrt_les_tau_051_against_oracle():
    # 200 steps: the lid's shear layer has spread over ~10 nodes, the LES
    # viscosity is active along it (tau = 0.51 is close to the stability limit)
    fp32_parity(cfg(nx=128, ny=128, nz=128, enable_les=1, smagorinsky=0.17, tau=0.51), 200)


def test_cavity_256_mrt_les_against_oracle():
    fp32_parity(cfg(nx=256, ny=256, nz=256, enable_les=1, smagorinsky=0.17, tau=0.6), 20)


def test_channel_512x128x128_against_oracle():
    c = cfg(nx=512, ny=128, nz=128, casetype=nflow.CHANNEL, enable_les=1, smagorinsky=0.17, tau=0.56, inlet_peak=0.05,
            perturb=0.01, seed=2024)
    fp32_parity(c, 40)


def test_cavity_256_conserves_mass_and_mirror_symmetry():
    """The cavity conserves mass exactly in exact arithmetic and is symmetric
    under z -> nz-1-z (u_z -> -u_z); FP32 keeps both to round-off."""
    n = 256
    s = run_gpu(cfg(nx=n, ny=n, nz=n, enable_les=1, smagorinsky=0.17, tau=0.6), 1000)
    m, mom, _ = s.diagnostics()
    assert abs(m - n ** 3) / n ** 3 < 2e-6
    r, ux, uy, uz = s.read_macro()
    assert np.isfinite(ux).all()
    flip = ux[:, ::-1, :]
    assert np.linalg.norm(ux - flip) / np.linalg.norm(ux) < 1e-4
    assert np.linalg.norm(uz + uz[:, ::-1, :]) / np.linalg.norm(ux) < 1e-4
    assert abs(mom[2]) / abs(mom[0]) < 1e-4


@pytest.mark.parametrize("storage", [nflow.F32, nflow.F16])
def test_cavity_256_is_deterministic(storage):
    c = cfg(nx=256, ny=256, nz=256, enable_les=1, smagorinsky=0.17, tau=0.6)
    a = run_gpu(c, 120, storage=storage).read_macro()
    b = run_gpu(c, 120, storage=storage).read_macro()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("storage", [nflow.F32, nflow.F16])
def test_lattice_beyond_32_bit_element_indices_matches_a_short_box(storage):
    """512 x 1600 x 512 (419 M nodes): each population buffer holds 19 x 419 M
    elements, more than 2^32, so every kernel addresses it through windows
    (a 64-bit base per workgroup, 32-bit offsets).  The FP32 lattice runs a
    two-step launch (y chunks of 48) and a single step, the FP16 one three
    single steps.  From rest, three steps disturb only the four planes under
    the lid (the state moves one plane per step), at the far end of the
    buffers, so those must equal, bitwise, the top planes of an 8-plane box
    of the same x-z extent, and every other plane must equal that box's
    untouched bottom plane; the FP64 mass sums must differ by exactly the
    added rest planes."""
    nx, ny, nz = 512, 1600, 512
    assert 19 * nx * ny * nz > 2 ** 32
    # rounding to nearest: stochastic rounding seeds each node by its id, which
    # differs between the two boxes at the same place under the lid
    c = dict(nx=nx, nz=nz, enable_les=1, smagorinsky=0.17, tau=0.6, stochastic_rounding=0)
    steps = 3
    small = run_gpu(cfg(ny=8, **c), steps, storage=storage)
    ms = small.diagnostics()[0]
    sm = small.read_macro()
    node_mass = float(np.sum(small.read_arF()[1, 1, 1].astype(np.float64)))  # a rest node, as the FP64 sum sees it
    small.close()
    big = nflow.Batch([gpu_input(cfg(ny=ny, **c), storage)])
    assert big.step2_height() == (48 if storage == nflow.F32 else 0)
    big.initialize()
    big.run(steps)
    mb = big.diagnostics(0)[0]
    bm = big.read_macro(0)
    big.close()
    for a, b in zip(bm, sm):  # [ny][nz][nx]
        assert np.array_equal(a[-(steps + 1):], b[-(steps + 1):])
        assert (a[:-(steps + 1)] == b[0]).all()
    assert abs(mb - (ms + (ny - 8) * nx * nz * node_mass)) < 1e-9 * mb

