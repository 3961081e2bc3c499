"""Stochastic rounding of the 16-bit stores (nflow_input.stochastic_rounding):
the device conversion is the host rule of test_rounding_cpu bit for bit;
runs are deterministic, bitwise equal in a batch and alone and across a
checkpoint restart, within the 16-bit tolerance of the FP64 oracle, and on a
long low-viscosity run much closer to FP32 than round-to-nearest, whose
step-to-step correlated rounding error accumulates."""
import numpy as np
import pytest

import nflow
from helpers import TOL_U_F16, cfg, macro_errors, run_oracle
from test_rounding_cpu import sr_bf16, sr_f16

pytestmark = pytest.mark.gpu


def _values(rng, n):
    v = np.concatenate([rng.uniform(-0.3, 0.3, n), rng.uniform(-0.02, 0.02, n), rng.uniform(-7e-5, 7e-5, n // 4),
                        [0.0, -0.0, 2.0 ** -7, -(2.0 ** -7), 2.0 ** -14, 0.5, 1.0 / 18, -1.0 / 36]])
    p = np.float32(2.0 ** -7)
    return np.concatenate([v, [np.nextafter(p, np.float32(0)), np.nextafter(p, np.float32(1))]]).astype(np.float32)


def test_device_rounding_is_the_host_rule_bit_for_bit():
    rng = np.random.default_rng(11)
    v = _values(rng, 200_000)
    seeds = rng.integers(0, 1 << 32, v.size, dtype=np.uint64).astype(np.uint32)
    h16 = nflow.round16(v, nflow.F16, seeds)
    assert np.array_equal(h16, sr_f16(v, seeds).view(np.uint16))
    hb = nflow.round16(v, nflow.BF16, seeds)
    assert np.array_equal(hb, (sr_bf16(v, seeds).view(np.uint32) >> 16).astype(np.uint16))
    # to nearest: numpy's float16 cast (round half to even)
    assert np.array_equal(nflow.round16(v, nflow.F16), v.astype(np.float16).view(np.uint16))


def test_device_rounding_is_unbiased_over_all_seed_levels():
    vals = np.array([2.0 ** -7 * (1 + 2.0 ** -20), 0.0123456, -0.0555555, 0.1, 1.0 / 18 + 1e-7], dtype=np.float32)
    lev = (np.arange(8192, dtype=np.uint64) << np.uint64(19)).astype(np.uint32)
    for v in vals:
        r = nflow.round16(np.full(lev.shape, v), nflow.F16, lev).view(np.float16).astype(np.float64)
        assert r.mean() == np.float64(v), (v, r.mean() - np.float64(v))


def test_stochastic_rounding_restarts_exactly_from_a_checkpoint(tmp_path):
    c = cfg(nx=24, ny=20, nz=18, enable_les=1, smagorinsky=0.17, tau=0.53, init_lid_velocity=1)
    full = solve(c, nflow.F16, 1, 60)
    s = solve(c, nflow.F16, 1, 25)
    p = tmp_path / "sr.ckp"
    s.save_checkpoint(str(p))
    inp = nflow.defaults(**c, storage=nflow.F16)
    inp.stochastic_rounding = 1
    t = nflow.Solver(inp)
    t.initialize()
    t.load_checkpoint(str(p))
    t.run(35)
    a, b = t.read_arF(), full.read_arF()
    assert np.array_equal(a, b), float(np.abs(a - b).max())
    # a solver that rounds to nearest refuses the stochastic-rounding checkpoint
    v = nflow.defaults(**c, storage=nflow.F16)
    v.stochastic_rounding = 0
    u = nflow.Solver(v)
    u.initialize()
    with pytest.raises(nflow.NflowError, match="parameters"):
        u.load_checkpoint(str(p))
    for x in (full, s, t, u):
        x.close()


def solve(c, storage, sr, steps):
    inp = nflow.defaults(**c, storage=storage)
    inp.stochastic_rounding = sr
    s = nflow.Solver(inp)
    s.initialize()
    s.run(steps)
    return s


@pytest.mark.parametrize("storage", [nflow.F16, nflow.BF16])
def test_stochastic_rounding_is_deterministic_and_equal_in_a_batch(storage):
    cs = [cfg(nx=33, ny=17, nz=9, enable_les=1, smagorinsky=0.17, tau=0.52, init_lid_velocity=1),
          cfg(nx=40, ny=12, nz=10, casetype=nflow.CHANNEL, enable_les=1, smagorinsky=0.17, tau=0.56, inlet_peak=0.05,
              perturb=0.01, seed=7),
          cfg(nx=16, ny=12, nz=10, enable_mrt=0, tau=0.6, init_lid_velocity=1)]
    srs = [1, 1, 0]
    inputs = []
    for c, sr in zip(cs, srs):
        i = nflow.defaults(**c, storage=storage)
        i.stochastic_rounding = sr
        inputs.append(i)
    b = nflow.Batch(inputs)
    b.set_tuning(128, 1, 2, 4)
    b.initialize()
    b.run(41)
    for k, (c, sr) in enumerate(zip(cs, srs)):
        a1 = solve(c, storage, sr, 41).read_arF()
        a2 = solve(c, storage, sr, 41).read_arF()
        assert np.array_equal(a1, a2) and np.array_equal(b.read_arF(k), a1), k
    # the rounding mode matters: stochastic and nearest differ
    assert not np.array_equal(solve(cs[0], storage, 1, 41).read_arF(), solve(cs[0], storage, 0, 41).read_arF())
    b.close()


def test_stochastic_rounding_within_16_bit_tolerance_of_the_oracle():
    c = cfg(nx=32, ny=32, nz=32, enable_les=1, smagorinsky=0.17, tau=0.55)
    eu, _ = macro_errors(solve(c, nflow.F16, 1, 300).read_macro(), run_oracle(c, 300).read_macro())
    assert eu < TOL_U_F16, eu


def test_stochastic_rounding_bounds_long_run_drift():
    # a served cavity of the full-length request stream: 67^3, Re 768 (tau 0.5131), 18,998 steps
    e, re = 67, 768.0
    c = cfg(nx=e, ny=e, nz=e, enable_les=1, smagorinsky=0.17, tau=0.5 + 3 * 0.05 * e / re)
    steps = 18998
    u32 = np.stack(solve(c, nflow.F32, 0, steps).read_macro()[1:]).astype(np.float64)
    err = {}
    for sr in (0, 1):
        u16 = np.stack(solve(c, nflow.F16, sr, steps).read_macro()[1:]).astype(np.float64)
        err[sr] = np.linalg.norm(u16 - u32) / np.linalg.norm(u32)
    print(f"rel-L2 u vs FP32 after {steps} steps: nearest {err[0]:.2e}, stochastic {err[1]:.2e}")
    assert err[1] < 1e-3 and err[1] < err[0] / 4, err
