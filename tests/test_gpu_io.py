"""Checkpoint files and output writers on device data."""
import numpy as np
import pytest

import nflow
from helpers import cfg, run_gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage", [nflow.F32, nflow.F16])
def test_checkpoint_file_restart(tmp_path, storage):
    c = cfg(nx=20, ny=16, nz=12, enable_les=1, smagorinsky=0.17, casetype=1)
    full = run_gpu(c, 70, storage=storage)
    s = run_gpu(c, 30, storage=storage)
    p = tmp_path / "state.ckp"
    s.save_checkpoint(str(p))
    assert nflow.checkpoint_info(str(p)) == (20, 16, 12, 30)
    snap = s.read_arF()
    t = run_gpu(c, 0, storage=storage)
    t.load_checkpoint(str(p))
    assert t.steps_done() == 30
    if storage == nflow.F32:
        assert np.array_equal(t.read_arF(), snap)  # export straight after a load
    t.run(40)
    assert t.steps_done() == 70
    if storage == nflow.F32:
        assert np.array_equal(t.read_arF(), full.read_arF())
    else:  # the import re-rounds the streamed state to FP16
        a, b = t.read_macro()[1], full.read_macro()[1]
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-3


def test_checkpoint_rejects_other_extents(tmp_path):
    s = run_gpu(cfg(nx=10, ny=10, nz=10), 3)
    p = tmp_path / "a.ckp"
    s.save_checkpoint(str(p))
    t = run_gpu(cfg(nx=12, ny=10, nz=10), 0)
    with pytest.raises(nflow.NflowError):
        t.load_checkpoint(str(p))
    u = run_gpu(cfg(nx=10, ny=10, nz=10, casetype=nflow.CHANNEL), 0)  # same extents, channel
    with pytest.raises(nflow.NflowError, match="casetype"):
        u.load_checkpoint(str(p))


def test_checkpoint_rejects_other_parameters_and_storage(tmp_path):
    c = cfg(nx=10, ny=9, nz=8, enable_les=1, smagorinsky=0.17, tau=0.6)
    s = run_gpu(c, 4)
    p = tmp_path / "b.ckp"
    s.save_checkpoint(str(p))
    for k, v in (("tau", 0.55), ("smagorinsky", 0.1), ("enable_mrt", 0), ("ref_vel", (0.04, 0.0, 0.0))):
        t = run_gpu(dict(c, **{k: v}), 0)
        with pytest.raises(nflow.NflowError, match="parameters"):
            t.load_checkpoint(str(p))
        t.close()
    u = run_gpu(c, 0, storage=nflow.F16)
    with pytest.raises(nflow.NflowError, match="storage"):
        u.load_checkpoint(str(p))
    w = run_gpu(dict(c, maxitr=999, nwrite=7), 0)  # run length and output cadence do not matter
    w.load_checkpoint(str(p))
    assert w.steps_done() == 4


def test_vtu_of_device_fields(tmp_path):
    s = run_gpu(cfg(nx=9, ny=8, nz=7, init_lid_velocity=1), 5)
    r, ux, uy, uz = s.read_macro()
    t, _ = s.node_type()
    p = tmp_path / "cavity.vtu"
    nflow.write_vtu(str(p), r, ux, uy, uz, t)
    from test_io_cpu import parse_vtu

    npts, ncell, a = parse_vtu(p)
    assert npts == 9 * 8 * 7 and np.array_equal(a["Density"], r.ravel())
    assert a["Moving Wall"].sum() == (9 - 2) * (7 - 2)


@pytest.mark.parametrize("casetype", [nflow.CAVITY, nflow.CHANNEL])
def test_checkpoint_restart_with_two_steps_per_launch(tmp_path, casetype):
    """A run with two steps per launch saved after an odd and an even number of
    steps, restarted, and continued equals one uninterrupted run with one step
    per launch, bit for bit (the export after each run reads the state of its
    last, single-step launch)."""
    c = cfg(nx=256, ny=14, nz=11, enable_les=1, smagorinsky=0.17, tau=0.57, casetype=casetype,
            inlet_peak=0.05, perturb=0.01, seed=11)
    ref = nflow.Solver(nflow.defaults(**c, storage=nflow.F32))
    ref.set_tuning(256, 0, 0, 10, 0)
    ref.initialize()
    ref.run(40)
    want = ref.read_arF()
    ref.close()
    for first in (17, 20):
        s = nflow.Solver(nflow.defaults(**c, storage=nflow.F32))
        s.set_tuning(256, 0, 0, 10, 4)
        s.initialize()
        s.run(first)
        p = tmp_path / f"two_{first}.ckp"
        s.save_checkpoint(str(p))
        s.close()
        t = nflow.Solver(nflow.defaults(**c, storage=nflow.F32))
        t.set_tuning(256, 0, 0, 10, 4)
        t.initialize()
        t.load_checkpoint(str(p))
        t.run(40 - first)
        assert t.steps_done() == 40
        assert np.array_equal(t.read_arF(), want), first
        t.close()
