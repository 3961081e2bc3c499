"""HIP kernels (through the C ABI) against the FP64 oracle.

Each case runs the same configuration on the GPU and in the oracle and
compares rho, u and the exported arF.  Tolerances are in helpers.py.
"""
import numpy as np
import pytest

import nflow
import oracle
from helpers import TOL_U_F16, cfg, fp32_parity, macro_errors, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("init_lid", [0, 1])
def test_initial_state_matches_oracle(init_lid):
    c = cfg(nx=12, ny=10, nz=8, init_lid_velocity=init_lid, ref_vel=(0.1, 0.0, 0.0))
    s = run_gpu(c, 0)
    L = run_oracle(c, 0)
    a = s.read_arF()
    b = L.read_arF()
    assert np.abs(a - b).max() < 1e-7
    # ghosts are zero
    assert not a[0].any() and not a[-1].any() and not a[:, 0].any() and not a[:, :, -1].any()
    t, d = s.node_type()
    to, do = L.node_type()
    assert np.array_equal(t, to) and np.array_equal(d, do)


@pytest.mark.parametrize("model,les,tau", [("srt", 0, 0.6), ("mrt", 0, 0.6), ("mrt", 1, 0.6), ("mrt", 1, 0.51),
                                           ("srt", 1, 0.55)])
def test_cavity_matches_oracle(model, les, tau):
    c = cfg(nx=16, ny=16, nz=16, enable_mrt=int(model == "mrt"), enable_les=les,
            smagorinsky=0.17 if les else 0.0, tau=tau)
    steps = 300
    _, s, o64, o32 = fp32_parity(c, steps)
    a, b, b32 = s.read_arF(), o64.read_arF(), o32.read_arF()
    assert np.abs(a - b).max() <= 4 * np.abs(b32 - b).max() + 1e-7


def test_cavity_32_srt_1000_steps():
    """The CPU-plain case: 32^3 SRT cavity, tau 0.6, lid 0.05, rest start, 1000 steps."""
    c = cfg(nx=32, ny=32, nz=32, enable_mrt=0, enable_les=0, tau=0.6)
    fp32_parity(c, 1000)


def test_channel_matches_oracle():
    c = cfg(nx=40, ny=12, nz=10, casetype=nflow.CHANNEL, enable_mrt=1, enable_les=1, smagorinsky=0.17, tau=0.56,
            inlet_peak=0.05, perturb=0.01, seed=1234)
    fp32_parity(c, 400)


@pytest.mark.parametrize("steps", [0, 1, 7])
def test_diagnostics_match_oracle(steps):
    c = cfg(nx=14, ny=12, nz=10, enable_les=1, smagorinsky=0.17)
    s = run_gpu(c, steps)
    L = run_oracle(c, steps)
    L32 = run_oracle(c, steps, fp32=True)
    m, mom, d = s.diagnostics()
    mo, momo, do = L.diagnostics()
    m32, mom32, d32 = L32.diagnostics()
    print(f"mass {m} {mo} {m32}; mom {mom} {momo} {mom32}; drho {d} {do} {d32}")
    assert abs(m - mo) / mo < 1e-6
    assert np.all(np.abs(mom - momo) <= 4 * np.abs(mom32 - momo) + 1e-6)
    if steps == 0:
        assert abs(d - do) / do < 1e-6  # PrevDen starts at 0: the residual equals the mass
    else:
        assert abs(d - do) <= 4 * abs(d32 - do) + 1e-6 * do


def test_decomposition_does_not_change_results():
    c = cfg(nx=17, ny=13, nz=11, enable_les=1, smagorinsky=0.17)
    inp = nflow.defaults(**c)
    ref = None
    for dec in [(1, 1, 1), (3, 3, 3), (2, 1, 4), (1, 5, 1)]:
        sl = nflow.decomposition(inp, *dec)
        s = run_gpu(c, 50, slices=sl)
        a = s.read_arF()
        if ref is None:
            ref = a
        else:
            assert np.array_equal(a, ref), f"decomposition {dec} changed the result"


def test_runs_are_bitwise_deterministic():
    c = cfg(nx=24, ny=20, nz=18, enable_les=1, smagorinsky=0.17)
    a = run_gpu(c, 100).read_arF()
    b = run_gpu(c, 100).read_arF()
    assert np.array_equal(a, b)


def test_split_runs_equal_one_run():
    c = cfg(nx=20, ny=16, nz=12, enable_les=1, smagorinsky=0.17)
    a = run_gpu(c, 60).read_arF()
    s = run_gpu(c, 0)
    for k in (1, 2, 17, 40):
        s.run(k)
    assert s.steps_done() == 60
    assert np.array_equal(s.read_arF(), a)


def test_checkpoint_restart_is_exact():
    c = cfg(nx=20, ny=16, nz=12, enable_les=1, smagorinsky=0.17)
    a = run_gpu(c, 80).read_arF()
    s = run_gpu(c, 30)
    snap = s.read_arF()
    t = run_gpu(c, 0)
    t.write_arF(snap)
    t.run(50)
    assert np.array_equal(t.read_arF(), a)


def test_import_matches_oracle_import():
    c = cfg(nx=12, ny=12, nz=12, enable_les=1, smagorinsky=0.17)
    L = run_oracle(c, 25)
    snap = L.read_arF()
    s = run_gpu(c, 0)
    s.write_arF(snap.astype(np.float32))
    s.run(40)
    L2 = oracle.Lattice(oracle.default_params(**c))
    L2.initialize()
    L2.write_arF(snap)
    L2.run(40)
    eu, er = macro_errors(s.read_macro(), L2.read_macro())
    print(f"import: rel-L2 u {eu:.2e} rho-1 {er:.2e}")
    assert eu < 1e-5 and er < 1e-3


@pytest.mark.parametrize("storage", [nflow.F16, nflow.BF16])
def test_reduced_precision_storage(storage):
    c = cfg(nx=24, ny=24, nz=24, enable_les=1, smagorinsky=0.17, tau=0.6)
    steps = 500
    s = run_gpu(c, steps, storage=storage)
    L = run_oracle(c, steps)
    eu, er = macro_errors(s.read_macro(), L.read_macro())
    print(f"storage {storage}: rel-L2 u {eu:.2e} rho-1 {er:.2e}")
    if storage == nflow.F16:
        assert eu < TOL_U_F16
    else:
        assert np.isfinite(eu) and eu < 0.1  # BF16 is reported, not held to 1e-3


def test_bad_inputs_raise():
    with pytest.raises(nflow.NflowError):
        nflow.Solver(nflow.defaults(nx=8, ny=8, nz=8, tau=0.5))
    s = nflow.Solver(nflow.defaults(nx=8, ny=8, nz=8))
    with pytest.raises(nflow.NflowError):
        s.run(3)  # before initialize
    inp = nflow.defaults(nx=8, ny=8, nz=8)
    bad = np.array([[[0, 8], [0, 8], [0, 4]]], dtype=np.int32)  # does not cover the box
    with pytest.raises(nflow.NflowError):
        nflow.Solver(inp, bad)


def test_copy_ceilings_run():
    # the measurement helper in both modes: equal arrays and the FP32 population layout
    for streams in (1, 19, 0):
        gbs, ms = nflow.copy_bandwidth(19 * 64 * 4 * 2048, streams=streams, reps=2)
        assert gbs > 0 and ms > 0, streams
    with pytest.raises(nflow.NflowError):
        nflow.copy_bandwidth(100, streams=0, reps=1)  # smaller than one 19-run block


def test_valu_probe_runs():
    g, ms = nflow.valu_rate(reps=2)
    assert g > 0 and ms > 0
    with pytest.raises(nflow.NflowError):
        nflow.valu_rate(reps=0)
