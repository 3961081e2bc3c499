"""Batched and served execution: every lattice gives the same result as its
solo run (bitwise for the same storage), whatever else shares the launches;
a small lattice of each mix is also checked against the FP64 oracle."""
import time

import numpy as np
import pytest

import nflow
import workloads
from helpers import TOL_U_F16, cfg, fp32_parity, macro_errors, run_oracle

pytestmark = pytest.mark.gpu


def solo_arF(inp, steps):
    s = nflow.Solver(inp)
    s.initialize()
    s.run(steps)
    a = s.read_arF()
    s.close()
    return a


def mixed_inputs(storage=nflow.F32):
    return [
        nflow.defaults(**cfg(nx=16, ny=12, nz=10, enable_mrt=0, tau=0.6), storage=storage),
        nflow.defaults(**cfg(nx=33, ny=17, nz=9, enable_les=1, smagorinsky=0.17, tau=0.52), storage=storage),
        nflow.defaults(**cfg(nx=40, ny=12, nz=10, casetype=nflow.CHANNEL, enable_les=1, smagorinsky=0.17, tau=0.56,
                             inlet_peak=0.05, perturb=0.01, seed=99), storage=storage),
        nflow.defaults(**cfg(nx=7, ny=5, nz=3, enable_mrt=1, tau=0.8), storage=storage),
        nflow.defaults(**cfg(nx=48, ny=48, nz=48, enable_les=1, smagorinsky=0.17, tau=0.6, init_lid_velocity=1),
                       storage=storage),
    ]


@pytest.mark.parametrize("storage", [nflow.F32, nflow.F16])
@pytest.mark.parametrize("tuning", [(256, 0, 0, 0), (64, 1, 0, 4), (512, 0, 4, 10), (128, 1, 2, 2)])
def test_batch_equals_solo_runs(storage, tuning):
    inputs = mixed_inputs(storage)
    steps = 37
    b = nflow.Batch(inputs)
    b.set_tuning(*tuning)
    b.initialize()
    b.run(steps)
    for i, inp in enumerate(inputs):
        assert np.array_equal(b.read_arF(i), solo_arF(inp, steps)), (i, tuning)


def test_autotuned_batch_equals_solo_runs():
    inputs = mixed_inputs()
    b = nflow.Batch(inputs)
    t = b.autotune(steps_per_trial=4)
    assert t["block"] in (64, 128, 256, 512) and t["ms_per_step"] > 0
    b.run(25)  # autotune leaves every lattice re-initialised
    for i, inp in enumerate(inputs):
        assert np.array_equal(b.read_arF(i), solo_arF(inp, 25)), i


def test_batched_cavities_against_oracle():
    """Config-4 style mix (seeded sizes/Re, SRT or MRT-LES), FP16 storage; the
    smallest lattice is compared with the FP64 oracle, the rest with solo runs."""
    lat = workloads.batched_cavities(seed=3, count=6, storage=nflow.F16)
    for d in lat:  # keep the test small: shrink edges, keep Re/model draws
        e = 16 + d["edge"] // 8
        nu = 0.05 * e / d["re"]
        d["input"].nx = d["input"].ny = d["input"].nz = d["input"].h = e
        d["input"].tau = 3 * nu + 0.5
    inputs = [d["input"] for d in lat]
    b = nflow.Batch(inputs)
    b.autotune(steps_per_trial=4)
    steps = 200
    b.run(steps)
    k = int(np.argmax([i.tau for i in inputs]))  # the most stable lattice (SRT near tau = 0.5 may diverge)
    inp = inputs[k]
    c = dict(nx=inp.nx, ny=inp.ny, nz=inp.nz, tau=inp.tau, enable_mrt=inp.enable_mrt, enable_les=inp.enable_les,
             smagorinsky=inp.smagorinsky, ref_vel=(0.05, 0.0, 0.0), init_lid_velocity=0, casetype=1)
    L = run_oracle(c, steps)
    eu, _ = macro_errors(b.read_macro(k), L.read_macro())
    print(f"lattice {k} ({inp.nx}^3, mrt={inp.enable_mrt}, tau={inp.tau:.4f}): rel-L2 u {eu:.2e}")
    assert eu < TOL_U_F16
    for i, inp in enumerate(inputs):
        s = nflow.Solver(inp)
        s.initialize()
        s.run(steps)
        assert np.array_equal(b.read_macro(i)[1], s.read_macro()[1], equal_nan=True), i


@pytest.mark.parametrize("storage", [nflow.F16, nflow.F32, nflow.BF16])
def test_server_results_equal_solo_results(storage):
    inputs = mixed_inputs(storage)
    steps = [30, 5, 41, 1, 17]
    srv = nflow.Server(budget_bytes=512 << 20, storage=storage, max_resident=3, steps_per_tick=8)
    for i, (inp, n) in enumerate(zip(inputs, steps)):
        srv.submit(inp, n, tag=100 + i, want_fields=True)
        time.sleep(0.002 * i)
    srv.drain()
    res = {r.tag: r for r in srv.poll(16)}
    assert sorted(res) == [100 + i for i in range(5)]
    for i, (inp, n) in enumerate(zip(inputs, steps)):
        r = res[100 + i]
        assert r.status == 0 and r.steps == n and r.done_s >= r.admit_s >= r.submit_s
        shape = (inp.ny, inp.nz, inp.nx)
        got = srv.fetch_macro(100 + i, shape)
        s = nflow.Solver(inp)
        s.initialize()
        s.run(n)
        want = s.read_macro()
        for a, w in zip(got, want):
            assert np.array_equal(a, w), i
        m, mom, d = s.diagnostics()
        assert r.mass == m and list(r.mom) == list(mom)
    st = srv.stats()
    assert st["completed"] == 5 and st["peak_resident"] <= 3
    srv.close()


def test_server_budget_admission():
    small = nflow.defaults(**cfg(nx=16, ny=16, nz=16))
    big = nflow.defaults(**cfg(nx=200, ny=200, nz=200))  # ~2.4 GB at FP16, over budget
    srv = nflow.Server(budget_bytes=64 << 20, storage=nflow.F16, max_resident=8, steps_per_tick=4)
    srv.submit(small, 10, tag=1)
    srv.submit(big, 10, tag=2)
    srv.submit(small, 3, tag=3)
    srv.drain()
    res = {r.tag: r for r in srv.poll(8)}
    assert res[1].status == 0 and res[3].status == 0
    assert res[2].status == -5  # NFLOW_EBUDGET
    st = srv.stats()
    assert st["rejected"] == 1 and st["peak_bytes"] <= 64 << 20
    srv.close()


def test_server_backfills_behind_a_blocked_request():
    """A request that must wait for memory does not stop smaller ones that fit."""
    a = nflow.defaults(**cfg(nx=64, ny=64, nz=64))
    srv = nflow.Server(budget_bytes=30 << 20, storage=nflow.F16, max_resident=8, steps_per_tick=2)
    srv.submit(a, 400, tag=1)   # ~20 MB, admitted first
    srv.submit(a, 10, tag=2)    # does not fit until 1 leaves
    small = nflow.defaults(**cfg(nx=12, ny=12, nz=12))
    srv.submit(small, 10, tag=3)  # fits beside 1
    srv.drain()
    res = {r.tag: r for r in srv.poll(8)}
    assert all(r.status == 0 for r in res.values())
    assert res[3].admit_s < res[1].done_s  # backfilled while 1 was resident
    assert res[2].admit_s >= res[1].done_s - 1e-3
    srv.close()


def test_server_concurrent_submitters():
    """Four host threads submit while the worker runs and the main thread
    polls; every request completes with the mass its solo run gives (FP32,
    bitwise), and the counters add up."""
    import threading

    inputs = mixed_inputs(nflow.F32)
    plan = [(t, k, (t * 6 + k) % len(inputs), 3 + (t + 2 * k) % 9) for t in range(4) for k in range(6)]
    srv = nflow.Server(budget_bytes=256 << 20, storage=nflow.F32, max_resident=4, steps_per_tick=5)
    errors = []

    def submit(t):
        try:
            for tt, k, i, n in plan:
                if tt == t:
                    srv.submit(inputs[i], n, tag=1000 * (t + 1) + k)
                    time.sleep(0.001)
        except Exception as e:  # pragma: no cover - reported below
            errors.append(e)

    threads = [threading.Thread(target=submit, args=(t,)) for t in range(4)]
    for th in threads:
        th.start()
    res = {}
    deadline = time.time() + 60
    while len(res) < len(plan) and time.time() < deadline:
        for r in srv.poll(64, wait_ms=100):
            res[r.tag] = r
    for th in threads:
        th.join()
    assert not errors and len(res) == len(plan)
    solo = {}
    for t, k, i, n in plan:
        r = res[1000 * (t + 1) + k]
        assert r.status == 0 and r.steps == n
        if (i, n) not in solo:
            s = nflow.Solver(inputs[i])
            s.initialize()
            s.run(n)
            solo[(i, n)] = s.diagnostics()[0]
            s.close()
        assert r.mass == solo[(i, n)], (t, k)
    st = srv.stats()
    assert st["submitted"] == st["completed"] == len(plan) and st["peak_resident"] <= 4
    srv.close()


def test_server_close_abandons_queued_work_promptly():
    big = nflow.defaults(**cfg(nx=64, ny=64, nz=64))
    srv = nflow.Server(budget_bytes=64 << 20, storage=nflow.F16, max_resident=2, steps_per_tick=4)
    for i in range(6):
        srv.submit(big, 1_000_000, tag=i)
    time.sleep(0.2)
    t0 = time.time()
    srv.close()  # the worker finishes its current tick and exits
    assert time.time() - t0 < 10


@pytest.mark.parametrize("storage", [nflow.F16, nflow.BF16])
def test_server_tau_floor_keeps_low_tau_requests_in_fp32(storage):
    """Below the tau floor a request is stored in FP32 (its own batch) and its
    result equals the FP32 solo run; the others stay in the server's 16-bit
    storage and equal their 16-bit solo runs, both bitwise."""
    inputs = mixed_inputs(storage)
    steps = [19, 23, 31, 7, 11]
    srv = nflow.Server(budget_bytes=512 << 20, storage=storage, max_resident=4, steps_per_tick=6)
    srv.set_tau_floor(0.57)
    for i, (inp, n) in enumerate(zip(inputs, steps)):
        srv.submit(inp, n, tag=200 + i, want_fields=True)
    srv.drain()
    res = {r.tag: r for r in srv.poll(16)}
    low = [i for i, inp in enumerate(inputs) if inp.tau < 0.57]
    assert low and len(low) < len(inputs)
    for i, (inp, n) in enumerate(zip(inputs, steps)):
        assert res[200 + i].status == 0
        solo_in = nflow.defaults(**{k: getattr(inp, k) for k, _ in inp._fields_ if k != "ref_vel"},
                                 ref_vel=tuple(inp.ref_vel))
        solo_in.storage = nflow.F32 if i in low else storage
        s = nflow.Solver(solo_in)
        s.initialize()
        s.run(n)
        want = s.read_macro()
        s.close()
        got = srv.fetch_macro(200 + i, (inp.ny, inp.nz, inp.nx))
        for a, w in zip(got, want):
            assert np.array_equal(a, w), (i, i in low)
    srv.close()
    with pytest.raises(nflow.NflowError):
        nflow.Server(budget_bytes=64 << 20, storage=storage).set_tau_floor(-1.0)
