"""Pins of the oracle's tables and moment basis: closed forms and the
reference's literal tables (tests/golden/*.txt, extracted from main.cpp and
kernel.cu by tests/golden/make_fixtures.py)."""
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    return np.loadtxt(os.path.join(GOLDEN, name))


def test_velocity_set_matches_reference_literal():
    d, opp, w = oracle.tables()
    g = load("d3q19_tables.txt")
    # fixture rows are Direction[XDIM], Direction[ZDIM], Direction[YDIM], OppIndex
    assert np.array_equal(d[oracle_axis("x")], g[0])
    assert np.array_equal(d[oracle_axis("z")], g[1])
    assert np.array_equal(d[oracle_axis("y")], g[2])
    assert np.array_equal(opp, g[3].astype(int))


def oracle_axis(a):
    return {"x": 0, "z": 1, "y": 2}[a]  # XDIM = 0, ZDIM = 1, YDIM = 2 (common.h:13-15)


def test_d3q19_identities():
    d, opp, w = oracle.tables()
    e = np.stack([d[0], d[2], d[1]], axis=1)  # (q, [x, y, z])
    assert abs(w.sum() - 1.0) < 1e-15
    assert np.allclose(w @ e, 0.0, atol=1e-15)
    assert np.allclose(np.einsum("q,qa,qb->ab", w, e, e), np.eye(3) / 3.0, atol=1e-15)
    # fourth-order isotropy: sum w e_a e_b e_c e_d = (d_ab d_cd + d_ac d_bd + d_ad d_bc) / 9
    I = np.eye(3)
    iso = (np.einsum("ab,cd->abcd", I, I) + np.einsum("ac,bd->abcd", I, I) + np.einsum("ad,bc->abcd", I, I)) / 9.0
    assert np.allclose(np.einsum("q,qa,qb,qc,qd->abcd", w, e, e, e, e), iso, atol=1e-15)
    assert np.array_equal(e[opp], -e)
    assert sorted(set(np.round(w * 36).astype(int))) == [1, 2, 12]


def test_moment_matrix_equals_reference_literals():
    m = oracle.mrt_matrix()
    assert np.array_equal(m, load("mrt_matrix_main_cpp.txt"))
    assert np.array_equal(m, load("mrt_matrix_kernel_cu.txt"))
    assert np.count_nonzero(m) == 213


def test_moment_matrix_is_orthogonal_and_inverse_exact():
    m = oracle.mrt_matrix()
    g = m @ m.T
    norms = [19, 2394, 252, 10, 40, 10, 40, 10, 40, 36, 72, 12, 24, 4, 4, 4, 8, 8, 8]
    assert np.array_equal(g, np.diag(norms).astype(float))
    rc, minv = oracle.mrt_inverse_exact(m)
    assert rc == 0
    assert np.allclose(minv @ m, np.eye(19), atol=1e-14)
    assert np.allclose(minv, np.linalg.inv(m), atol=1e-14)
    # a non-orthogonal matrix is refused
    bad = m.copy()
    bad[1, 0] += 1.0
    assert oracle.mrt_inverse_exact(bad)[0] == -1


def test_equilibrium_moments_closed_form():
    """feq reproduces rho, rho u and rho/3 I + rho u u (A.2)."""
    d, _, _ = oracle.tables()
    e = np.stack([d[0], d[2], d[1]], axis=1)
    rng = np.random.default_rng(1)
    for _ in range(20):
        rho = 0.9 + 0.2 * rng.random()
        u = 0.1 * (rng.random(3) - 0.5)
        f = oracle.cell_feq(rho, u)
        assert abs(f.sum() - rho) < 1e-14
        assert np.allclose(f @ e, rho * u, atol=1e-15)
        assert np.allclose(np.einsum("q,qa,qb->ab", f, e, e), rho / 3 * np.eye(3) + rho * np.outer(u, u), atol=1e-15)


def test_inverse_against_reference_inverseMRT():
    """inverseMRT.cpp built from the reference sources (oracle/_ref)."""
    R = oracle.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    import ctypes

    m = oracle.mrt_matrix().astype(np.float32)
    out = np.zeros((19, 19), dtype=np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    R.ref_inverse_mrt(m.ctypes.data_as(fp), out.ctypes.data_as(fp))
    _, exact = oracle.mrt_inverse_exact(m.astype(np.float64))
    assert np.abs(out - exact).max() < 2e-6
