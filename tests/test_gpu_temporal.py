"""Two time steps per launch (nflow_tuning.temporal, k_step2; FP32 cavity and
channel lattices) against one step per launch: the same state bit for bit,
for both collision models, partial tiles and y chunks, several x tiles, the
channel's inlet and outlet in one or two tiles, tiny boxes, the benchmark
size, split runs, a read-back after every run, and read-backs after runs that
end with a two-step launch (the read-back streams the previous step's state,
which such a launch does not leave in memory: it is re-run as two single
steps first)."""
import numpy as np
import pytest

import nflow
from helpers import cfg

pytestmark = pytest.mark.gpu


def run(c, storage, sr, steps_list, temporal):
    inp = nflow.defaults(**c, storage=storage)
    inp.stochastic_rounding = sr
    b = nflow.Batch([inp])
    b.set_tuning(256, 0, 0, 0, temporal)
    b.initialize()
    out = []
    for n in steps_list:
        b.run(n)
        out.append((b.read_arF(0), b.read_macro(0), b.diagnostics(0)))
    b.close()
    return out


CASES = [
    cfg(nx=64, ny=48, nz=40, enable_les=1, smagorinsky=0.17, tau=0.56, init_lid_velocity=1),
    cfg(nx=37, ny=29, nz=23, enable_mrt=0, tau=0.6, init_lid_velocity=1),
    cfg(nx=70, ny=5, nz=17, enable_les=1, smagorinsky=0.17, tau=0.53),
    cfg(nx=3, ny=2, nz=5, enable_les=1, smagorinsky=0.17, tau=0.7, init_lid_velocity=1),
    # wider than one x tile: halo columns between tiles, a partial last tile
    cfg(nx=300, ny=11, nz=20, enable_les=1, smagorinsky=0.17, tau=0.55, init_lid_velocity=1),
    # two full x tiles, a partial z tile
    cfg(nx=512, ny=7, nz=21, enable_les=1, smagorinsky=0.17, tau=0.58, init_lid_velocity=1),
    # channels: inlet and outlet in one tile; the outlet in a second tile; the
    # outlet alone in the second tile (u(nx - 2) from the halo column)
    cfg(nx=40, ny=12, nz=10, casetype=nflow.CHANNEL, enable_les=1, smagorinsky=0.17, tau=0.56, inlet_peak=0.05,
        perturb=0.01, seed=3),
    cfg(nx=300, ny=9, nz=18, casetype=nflow.CHANNEL, enable_les=1, smagorinsky=0.17, tau=0.56, inlet_peak=0.05,
        perturb=0.01, seed=5),
    cfg(nx=257, ny=10, nz=11, casetype=nflow.CHANNEL, enable_mrt=0, tau=0.6, inlet_peak=0.05, perturb=0.01, seed=7),
]


@pytest.mark.parametrize("k", range(len(CASES)))
def test_two_steps_per_launch_equal_one(k):
    c = CASES[k]
    steps = [7, 8, 3, 20]
    ref = run(c, nflow.F32, 0, steps, 0)
    for h in (1, 3, 8, 12, 24, 32, 64):  # 8-64: the heights the automatic rule picks
        got = run(c, nflow.F32, 0, steps, h)
        for j, ((a, m, d), (a0, m0, d0)) in enumerate(zip(got, ref)):
            assert np.array_equal(a, a0), (h, j)
            for x, y in zip(m, m0):
                assert np.array_equal(x, y), (h, j)
            assert d[0] == d0[0] and np.array_equal(d[1], d0[1]) and d[2] == d0[2], (h, j)


def test_two_steps_per_launch_at_256_cubed():
    c = cfg(nx=256, ny=256, nz=256, enable_les=1, smagorinsky=0.17, tau=0.6)
    out = []
    for h in (0, 32):
        b = nflow.Batch([nflow.defaults(**c, storage=nflow.F32)])
        b.set_tuning(512, 0, 0, 0, h)
        b.initialize()
        b.run(9)
        out.append((b.read_macro(0), b.diagnostics(0)))
        b.close()
    (m, d), (m2, d2) = out
    for x, y in zip(m, m2):
        assert np.array_equal(x, y)
    assert d[0] == d2[0] and np.array_equal(d[1], d2[1]) and d[2] == d2[2]


def test_two_steps_per_launch_on_the_512x128x128_channel():
    """The channel of the precision runs at its automatic chunk height (16):
    two x tiles, inlet and outlet in different tiles, 8 y chunks."""
    c = cfg(nx=512, ny=128, nz=128, casetype=nflow.CHANNEL, enable_les=1, smagorinsky=0.17, tau=0.56,
            inlet_peak=0.05, perturb=0.01, seed=11)
    out = []
    for h in (0, -1):
        b = nflow.Batch([nflow.defaults(**c, storage=nflow.F32)])
        b.set_tuning(512, 0, 0, 0, h)
        assert b.step2_height() == (0 if h == 0 else 16)
        b.initialize()
        got = []
        for n in (9, 6):
            b.run(n)
            got.append((b.read_macro(0), b.diagnostics(0)))
        b.close()
        out.append(got)
    for (m, d), (m2, d2) in zip(*out):
        for x, y in zip(m, m2):
            assert np.array_equal(x, y)
        assert d[0] == d2[0] and np.array_equal(d[1], d2[1]) and d[2] == d2[2]


@pytest.mark.parametrize("storage,sr", [(nflow.F16, 1), (nflow.BF16, 0)])
def test_temporal_is_ignored_for_16_bit_storage(storage, sr):
    a = run(CASES[0], storage, sr, [9], 0)[0][0]
    b = run(CASES[0], storage, sr, [9], 16)[0][0]
    assert np.array_equal(a, b)


def test_temporal_is_ignored_where_it_does_not_apply():
    # a pipe (solid nodes) and a batch of two lattices keep one step per launch
    pipe = cfg(nx=30, ny=20, nz=20, casetype=nflow.PIPE, pipe_radius=8, pipe_cy=10, pipe_cz=10, tau=0.6)
    a = run(pipe, nflow.F32, 0, [9], 0)[0][0]
    b = run(pipe, nflow.F32, 0, [9], 16)[0][0]
    assert np.array_equal(a, b)
    two = [nflow.defaults(**c, storage=nflow.F32) for c in CASES[:2]]
    out = []
    for h in (0, 16):
        bt = nflow.Batch(two)
        bt.set_tuning(256, 0, 0, 0, h)
        assert bt.step2_height() == 0
        bt.initialize()
        bt.run(9)
        out.append([bt.read_arF(i) for i in range(2)])
        bt.close()
    for x, y in zip(*out):
        assert np.array_equal(x, y)


def test_automatic_chunk_height():
    """temporal = -1 (the default): the chunk height (8-64 planes) of the least
    estimated launch time, two steps per launch only where the estimate beats
    40 G node updates per second (256 CUs).  Each case's height is the fastest
    measured among 8, 12, 16, 24, 32, 48 and 64, and each case left at one
    step per launch measured no faster with two (within 1%; 128^3 7.6%)."""
    def height(nx, ny, nz, storage=nflow.F32, casetype=nflow.CAVITY, temporal=-1):
        b = nflow.Batch([nflow.defaults(**cfg(nx=nx, ny=ny, nz=nz, casetype=casetype), storage=storage)])
        b.set_tuning(512, 0, 0, 0, temporal)
        h = b.step2_height()
        b.close()
        return h
    assert nflow.device_info()[1] == 256
    assert height(256, 256, 256) == 32  # 32 tiles x 8 chunks
    assert height(256, 128, 256) == 16
    assert height(512, 256, 512) == 64  # 128 tiles x 4 chunks
    assert height(256, 64, 256) == 8
    assert height(256, 96, 256) == 12
    assert height(192, 192, 192) == 24  # partial x tiles
    assert height(384, 384, 384) == 48
    assert height(200, 300, 100) == 16
    assert height(128, 128, 128) == 0  # half the lanes idle
    assert height(160, 160, 160) == 0
    assert height(288, 288, 288) == 0  # second x tile an eighth full
    assert height(512, 64, 64) == 0  # too few tiles for the CUs
    assert height(48, 48, 48) == 0
    assert height(256, 256, 256, storage=nflow.F16) == 0
    assert height(512, 128, 128, casetype=nflow.CHANNEL) == 16  # 32 tiles x 8 chunks
    assert height(768, 128, 128, casetype=nflow.CHANNEL) == 8
    assert height(256, 128, 128, casetype=nflow.CHANNEL) == 8
    assert height(320, 128, 128, casetype=nflow.CHANNEL) == 0
    assert height(512, 128, 128, casetype=nflow.PIPE) == 0
    assert height(256, 256, 256, temporal=24) == 24
    assert height(256, 256, 256, temporal=0) == 0


@pytest.mark.parametrize("k", [0, 7])
def test_read_back_after_runs_that_end_with_a_pair(k, tmp_path):
    """Runs of even length with no read-back between end with a two-step
    launch, after which the other buffer holds the state two steps back; each
    kind of read-back (macro fields, populations, diagnostics, the
    allocation-free diagnostics, a checkpoint, a change of tuning) first
    re-runs that launch as two single steps and must find the state of one
    step per launch, bit for bit, and so must the runs after it."""
    c = CASES[k]

    def sequence(temporal):
        s = nflow.Solver(nflow.defaults(**c, storage=nflow.F32))
        s.set_tuning(256, 0, 0, 0, temporal)
        s.initialize()
        out = []
        s.run(10)
        s.run(6)  # ends with a pair (h > 0)
        out.append(("macro", s.read_macro()))
        s.run(4)  # after a read-back: ends with single steps
        out.append(("arF", s.read_arF()))
        s.run(2)
        s.run(2)
        out.append(("noalloc", s.diagnostics_noalloc()))
        s.run(8)
        out.append(("diag", s.diagnostics()))
        s.run(12)
        s.run(2)
        path = str(tmp_path / f"t{temporal}.ckp")
        s.save_checkpoint(path)
        with open(path, "rb") as f:
            out.append(("checkpoint", f.read()))
        s.run(4)  # after a read-back: ends with single steps
        s.run(2)  # ends with a pair
        s.set_tuning(256, 0, 0, 0, 0)  # settles the pair before the tuning changes
        s.run(3)
        out.append(("after retune", s.read_arF()))
        out.append(("steps", s.steps_done()))
        s.close()
        return out

    ref = sequence(0)
    for h in (1, 32):
        got = sequence(h)
        for (name, a), (_, b) in zip(got, ref):
            if isinstance(a, tuple):
                for x, y in zip(a, b):
                    assert np.array_equal(np.asarray(x), np.asarray(y)), (h, name)
            elif isinstance(a, np.ndarray):
                assert np.array_equal(a, b), (h, name)
            else:
                assert a == b, (h, name)
