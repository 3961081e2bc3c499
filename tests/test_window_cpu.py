"""Address arithmetic of the step kernel's windowed buffer access (FP16/BF16
storage): for every node of every work item, each pulled donor element and
each stored element, addressed relative to the work item's window, must be
the element whole-array indexing names, lie inside the window and fit a
32-bit byte offset.  Pure index arithmetic, no device code."""
import numpy as np
import pytest

Q = 19
EX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
EY = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1])
EZ = np.array([0, 0, 0, -1, 1, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0, -1, 1, -1, 1])
OPP = np.array([0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15])


def pidx(q, n, b):
    """AoSoA element index: blocks of b nodes, 19 runs of b elements each."""
    return ((n // b) * Q + q) * b + (n & (b - 1))


def window_base(start, nx, nz, b):
    reach = nx * nz + nx  # |offset| of the farthest donor (e = (0, +-1, +-1))
    return (start - reach if start > reach else 0) & ~(b - 1)


@pytest.mark.parametrize("seed", range(6))
def test_window_addresses_match_whole_array_indexing(seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = (int(v) for v in rng.integers([1, 1, 1], [45, 10, 30]))
    n_nodes = nx * ny * nz
    pstride = (n_nodes + 127) // 128 * 128
    nodes = np.arange(n_nodes)
    x, z, y = nodes % nx, (nodes // nx) % nz, nodes // (nx * nz)
    for esize in (2, 4):
        b = 256 // esize
        total = pstride * Q
        for block in (64, 128, 256, 512):
            for start in range(0, n_nodes, block):
                sl = slice(start, min(start + block, n_nodes))
                n = nodes[sl]
                bn = window_base(start, nx, nz, b)
                skip = bn * Q
                nbytes = (total - skip) * esize
                nr = n - bn
                for q in range(Q):
                    sx, sy, sz = x[sl] - EX[q], y[sl] - EY[q], z[sl] - EZ[q]
                    inside = (sx >= 0) & (sx < nx) & (sy >= 0) & (sy < ny) & (sz >= 0) & (sz < nz)
                    off = EX[q] + EZ[q] * nx + EY[q] * nx * nz
                    rel = np.where(inside, pidx(q, nr - off, b), pidx(OPP[q], nr, b))
                    whole = np.where(inside, pidx(q, n - off, b), pidx(OPP[q], n, b))
                    assert (rel >= 0).all() and (rel + skip == whole).all()
                    assert (rel * esize < nbytes).all() and (rel * esize < 2 ** 32).all()
                    store = pidx(q, nr, b)
                    assert (store + skip == pidx(q, n, b)).all() and (store * esize < nbytes).all()


def test_window_stays_small_for_the_largest_accepted_plane():
    # validate_input accepts an x-z plane while (2 (nx nz + nx) + 1024) 19 * 4 < 2^32
    nx = nz = 3700
    plane = nx * nz + nx
    assert (2 * plane + 1024) * 19 * 4 < 2 ** 32
    b = 64
    start = 5 * plane
    bn = window_base(start, nx, nz, b)
    far = start + 511 + plane  # farthest donor of a 512-node item
    assert (pidx(Q - 1, far - bn, b) + 1) * 4 < 2 ** 32
