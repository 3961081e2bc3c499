"""Solid nodes (NODE_STRUCT) on the device: the pipe case and user masks,
against the FP64 oracle, plus the masked kernel's identity with the plain one
where nothing is solid."""
import numpy as np
import pytest

import nflow
from helpers import TOL_U_F16, cfg, fp32_parity, macro_errors, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


def pipe_cfg(**kw):
    c = cfg(nx=30, ny=20, nz=19, casetype=nflow.PIPE, pipe_radius=8.3, pipe_cy=9.5, pipe_cz=9.0, tau=0.7,
            ref_vel=(0.05, 0.0, 0.0), enable_mrt=1, enable_les=1, smagorinsky=0.17)
    c.update(kw)
    return c


def obstacle(shape, seed=4):
    rng = np.random.default_rng(seed)
    m = np.zeros(shape, dtype=bool)
    ny, nz, nx = shape
    m[ny // 3:ny // 2, nz // 4:3 * nz // 4, nx // 3:nx // 2] = True  # a block
    m |= rng.random(shape) < 0.04  # scattered single solid nodes
    return m


@pytest.mark.parametrize("model", [0, 1])
def test_pipe_matches_oracle(model):
    c = pipe_cfg(enable_mrt=model)
    res, s, o64, _ = fp32_parity(c, 150)
    t, d = s.node_type()
    to, do = o64.node_type()
    assert np.array_equal(t, to) and np.array_equal(d, do)
    r, ux, _, _ = s.read_macro()
    solid = t == nflow.NODE_STRUCT
    assert solid.any() and not r[solid].any() and not ux[solid].any()


@pytest.mark.parametrize("case", [1, 0])
def test_masked_lattice_matches_oracle(case):
    c = cfg(nx=24, ny=18, nz=14, casetype=case, enable_les=1, smagorinsky=0.17, tau=0.6, init_lid_velocity=1,
            inlet_peak=0.05, perturb=0.01, seed=5)
    mask = obstacle((18, 14, 24))
    res, s, o64, _ = fp32_parity(c, 120, mask=mask)
    m, mom, _ = s.diagnostics()
    mo, momo, _ = o64.diagnostics()
    assert abs(m - mo) < 1e-4 * mo
    a = s.read_arF()[1:-1, 1:-1, 1:-1]
    assert not a[mask].any()


def test_masked_kernel_with_nothing_solid_is_bitwise_plain():
    c = cfg(nx=20, ny=16, nz=12, enable_les=1, smagorinsky=0.17, init_lid_velocity=1)
    plain = run_gpu(c, 40).read_arF()
    masked = run_gpu(c, 40, mask=np.zeros((16, 12, 20), dtype=bool)).read_arF()
    assert np.array_equal(plain, masked)


def test_cavity_obstacle_conserves_mass():
    c = cfg(nx=24, ny=24, nz=24, enable_les=1, smagorinsky=0.17, ref_vel=(0.08, 0.0, 0.0))
    mask = np.zeros((24, 24, 24), dtype=bool)
    mask[8:16, 6:18, 10:14] = True
    s = run_gpu(c, 500, mask=mask)
    m, mom, _ = s.diagnostics()
    assert abs(m - (24 ** 3 - mask.sum())) < 2e-3  # FP32 round-off over 500 steps
    assert abs(mom[2]) < 1e-4


def test_pipe_reduced_precision():
    c = pipe_cfg(enable_les=0)
    ref = run_gpu(c, 300).read_macro()
    for storage, tol in ((nflow.F16, TOL_U_F16), (nflow.BF16, 1e-2)):
        eu, _ = macro_errors(run_gpu(c, 300, storage=storage).read_macro(), ref)
        assert eu < tol, (storage, eu)


def test_pipe_in_a_batch_and_on_the_server():
    inputs = [nflow.defaults(**pipe_cfg()), nflow.defaults(**cfg(nx=18, ny=14, nz=10, init_lid_velocity=1))]
    solo = []
    for inp in inputs:
        s = nflow.Solver(inp)
        s.initialize()
        s.run(33)
        solo.append(s.read_arF())
    b = nflow.Batch(inputs)  # the cavity runs in the masked instantiation here
    b.set_tuning(128, 1, 2, 2)
    b.initialize()
    b.run(33)
    for i in range(2):
        assert np.array_equal(b.read_arF(i), solo[i]), i
    srv = nflow.Server(budget_bytes=256 << 20, steps_per_tick=10)
    srv.submit(inputs[0], 33, tag=7, want_fields=True)
    srv.drain()
    done = srv.poll(4)
    assert [r.tag for r in done] == [7] and done[0].status == 0
    r, ux, _, _ = srv.fetch_macro(7, (20, 19, 30))
    s = nflow.Solver(inputs[0])
    s.initialize()
    s.run(33)
    assert np.array_equal(ux, s.read_macro()[1])
    srv.close()


def test_mask_after_initialize_is_an_error():
    s = run_gpu(cfg(), 0)
    with pytest.raises(nflow.NflowError):
        s.set_solid_mask(np.zeros((16, 16, 16), dtype=bool))
    with pytest.raises(ValueError):
        nflow.Solver(nflow.defaults(**cfg())).set_solid_mask(np.zeros((3, 3, 3)))
