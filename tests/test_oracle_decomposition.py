"""Decomposition API: oracle and product library against the reference's own
decomposition.cpp (built from the reference sources into oracle/_ref), plus
brute-force partition checks."""
import ctypes

import numpy as np
import pytest

import nflow
import oracle

SIZES = [(10, 10, 10), (1, 1, 1), (2, 5, 3), (17, 13, 11), (256, 256, 256), (512, 128, 128), (4, 3, 2), (7, 8, 9)]


def ref_decomp(nx, ny, nz):
    R = oracle.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    s = np.zeros((1471, 3, 2), dtype=np.int32)
    nt = R.ref_decomposition(nx, ny, nz, s.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), 1471)
    return s[:nt].copy()


@pytest.mark.parametrize("n", SIZES)
def test_decomposition_matches_reference(n):
    R = oracle.ref()
    if R is None:
        pytest.skip("oracle/_ref not built")
    xd, yd, zd = (ctypes.c_int(), ctypes.c_int(), ctypes.c_int())
    R.ref_decomp_counts(ctypes.byref(xd), ctypes.byref(yd), ctypes.byref(zd))
    ref = ref_decomp(*n)
    ours = oracle.decomposition(*n, xd.value, yd.value, zd.value)
    lib = nflow.decomposition(nflow.defaults(nx=n[0], ny=n[1], nz=n[2]), xd.value, yd.value, zd.value)
    assert np.array_equal(ours, ref)
    assert np.array_equal(lib, ref)


@pytest.mark.parametrize("n", SIZES)
def test_move_and_slice_chunk_match_reference(n):
    R = oracle.ref()
    if R is None:
        pytest.skip("oracle/_ref not built")
    ref = ref_decomp(*n)
    ip = ctypes.POINTER(ctypes.c_int)
    mv_ref = np.zeros((len(ref), 19), dtype=np.int32)
    R.ref_set_move_decomp(len(ref), ref.ctypes.data_as(ip), mv_ref.ctypes.data_as(ip))
    assert np.array_equal(oracle.set_move_decomp(ref), mv_ref)
    assert np.array_equal(nflow.set_move_decomp(ref), mv_ref)
    for t in range(len(ref)):
        out = np.zeros(10, dtype=np.int32)
        rc_ref = R.ref_slice_chunk(t, ref.ctypes.data_as(ip), out.ctypes.data_as(ip))
        rc_o, o = oracle.slice_chunk(t, ref)
        rc_n, nl = nflow.slice_chunk(t, ref)
        assert rc_ref == rc_o == rc_n
        assert np.array_equal(out, o) and np.array_equal(out, nl)


@pytest.mark.parametrize("dec", [(1, 1, 1), (3, 3, 3), (2, 5, 1), (4, 1, 7), (5, 5, 5)])
@pytest.mark.parametrize("n", [(10, 10, 10), (3, 7, 4), (9, 2, 16)])
def test_tiles_partition_the_box(n, dec):
    s = oracle.decomposition(*n, *dec)
    assert np.array_equal(nflow.decomposition(nflow.defaults(nx=n[0], ny=n[1], nz=n[2]), *dec), s)
    cover = np.zeros(n[::-1], dtype=int)  # [z][y][x] order irrelevant here
    for t in s:
        (x0, x1), (y0, y1), (z0, z1) = t
        cover[z0:z1, y0:y1, x0:x1] += 1
    assert (cover == 1).all()


def test_move_offsets_brute_force():
    """Move[t][q] points at node x + e_q inside the tile's ghost-padded block."""
    d, _, _ = oracle.tables()
    s = oracle.decomposition(7, 6, 5, 2, 3, 1)
    mv = oracle.set_move_decomp(s)
    for t, ((x0, x1), (y0, y1), (z0, z1)) in enumerate(s):
        nx2, nz2 = x1 - x0 + 2, z1 - z0 + 2
        for q in range(19):
            ex, ez, ey = int(d[0][q]), int(d[1][q]), int(d[2][q])
            # local node (x,y,z) sits at padded (x+1, y+1, z+1); the target is (x+1+ex, ...)
            for (x, y, z) in [(0, 0, 0), (x1 - x0 - 1, y1 - y0 - 1, z1 - z0 - 1)]:
                target = (x + 1 + ex) + (z + 1 + ez) * nx2 + (y + 1 + ey) * nx2 * nz2
                assert x + z * nx2 + y * nx2 * nz2 + mv[t][q] == target
