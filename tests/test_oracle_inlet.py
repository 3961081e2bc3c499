"""Known-answer pins of the oracle's channel inlet (orc_splitmix64,
orc_inlet_xi, orc_inlet_profile in oracle/lbm_oracle.c).

The reference's inlet is the uniform RefVel of kernel.cu:599-605; the duct
inlet of the channel case is u_x = peak * 16 eta(1-eta) zeta(1-zeta) *
(1 + perturb * xi), eta = (y + 0.5)/ny, zeta = (z + 0.5)/nz, with xi in
[-1, 1] fixed per inlet node by splitmix64 of (seed, y, z).  These tests pin
each piece against something the oracle's code does not contain: the
published splitmix64 output sequence, the closed-form parabola, and the
statistics a uniform variate must have.
"""
import numpy as np
import pytest

import oracle

GAMMA = 0x9E3779B97F4A7C15
M64 = (1 << 64) - 1
# splitmix64 seeded with 0: the first outputs of next() (Steele, Lea & Flood,
# "Fast splittable pseudorandom number generators", OOPSLA 2014; Vigna's
# reference splitmix64.c), each next() adding GAMMA to the state before mixing
SEED0_OUTPUTS = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC,
                 0x1B39896A51A8749B]


def xi_of(h):
    return 2.0 * ((h >> 11) * 2.0 ** -53) - 1.0


def test_splitmix64_published_sequence():
    # orc_splitmix64(x) mixes x + GAMMA, i.e. it is next() on state x
    for k, want in enumerate(SEED0_OUTPUTS):
        assert oracle.lib().orc_splitmix64((k * GAMMA) & M64) == want, k


def test_inlet_xi_known_answers():
    """xi(seed, y, z) = 2 * (splitmix64(seed ^ (y << 32 | z)) >> 11) / 2^53 - 1:
    choosing seed so that the key is k * GAMMA turns each published output
    into a known answer, which pins the packing of (y, z) into the key."""
    L = oracle.lib()
    assert L.orc_inlet_xi(0, 0, 0) == xi_of(SEED0_OUTPUTS[0])
    for k, (y, z) in enumerate([(0, 0), (3, 5), (127, 1), (0, 126), (64, 64)]):
        seed = ((k * GAMMA) & M64) ^ ((y << 32) | z)
        assert L.orc_inlet_xi(seed, y, z) == xi_of(SEED0_OUTPUTS[k]), (k, y, z)


def test_inlet_parabola_without_perturbation():
    """perturb = 0 gives the closed-form duct parabola, whose maximum over the
    continuum cross-section is the peak (at eta = zeta = 1/2)."""
    for ny, nz, peak in [(128, 128, 0.05), (7, 12, 0.08), (1, 3, 0.03)]:
        p = oracle.default_params(nx=8, ny=ny, nz=nz, casetype=oracle.CHANNEL, inlet_peak=peak, perturb=0.0,
                                  seed=12345)
        got = oracle.inlet_profile(p)
        eta = (np.arange(ny) + 0.5) / ny
        zeta = (np.arange(nz) + 0.5) / nz
        want = peak * 16.0 * np.outer(eta * (1 - eta), zeta * (1 - zeta))
        np.testing.assert_allclose(got, want, rtol=1e-15, atol=0.0)
        assert got.max() <= peak and got.min() > 0.0
    # odd extents put a node on the centre line: the peak itself
    p = oracle.default_params(nx=8, ny=9, nz=9, casetype=oracle.CHANNEL, inlet_peak=0.05, perturb=0.0, seed=1)
    assert oracle.inlet_profile(p)[4, 4] == pytest.approx(0.05, rel=1e-15)


def test_inlet_perturbation_is_a_function_of_seed_y_z_only():
    """The relative perturbation (u / parabola - 1) / perturb equals xi(seed,
    y, z) at every inlet node, lies in [-1, 1], and does not depend on the
    lattice extents, the peak or the perturbation amplitude."""
    L = oracle.lib()
    seed = 2024
    ref = None
    for nx, ny, nz, peak, amp in [(512, 128, 128, 0.05, 0.01), (16, 128, 128, 0.02, 0.3), (3, 40, 90, 0.05, 0.01)]:
        p = oracle.default_params(nx=nx, ny=ny, nz=nz, casetype=oracle.CHANNEL, inlet_peak=peak, perturb=amp,
                                  seed=seed)
        base = oracle.default_params(nx=nx, ny=ny, nz=nz, casetype=oracle.CHANNEL, inlet_peak=peak, perturb=0.0,
                                     seed=seed)
        rel = (oracle.inlet_profile(p) / oracle.inlet_profile(base) - 1.0) / amp
        xi = np.array([[L.orc_inlet_xi(seed, y, z) for z in range(nz)] for y in range(ny)])
        np.testing.assert_allclose(rel, xi, rtol=0, atol=1e-12)
        assert np.all(np.abs(xi) <= 1.0)
        if ref is None:
            ref = xi
        else:
            np.testing.assert_array_equal(xi, ref[:ny, :nz])


def test_inlet_xi_is_uniform_and_seed_dependent():
    """Over a 128 x 128 inlet, xi has the moments of U[-1, 1] and passes a
    Kolmogorov-Smirnov test; another seed gives uncorrelated values."""
    from scipy import stats

    L = oracle.lib()
    xs = np.array([[L.orc_inlet_xi(7, y, z) for z in range(128)] for y in range(128)]).ravel()
    n = xs.size
    assert abs(xs.mean()) < 4 * np.sqrt(1 / 3 / n)
    assert abs(xs.var() - 1 / 3) < 4 * np.sqrt(4 / 45 / n)  # Var(x^2) = 1/5 - 1/9 for U[-1, 1]
    assert stats.kstest(xs, stats.uniform(loc=-1, scale=2).cdf).pvalue > 1e-3
    ys = np.array([[L.orc_inlet_xi(8, y, z) for z in range(128)] for y in range(128)]).ravel()
    assert abs(np.corrcoef(xs, ys)[0, 1]) < 4 / np.sqrt(n)
    # neighbouring nodes are not correlated either
    g = xs.reshape(128, 128)
    assert abs(np.corrcoef(g[:, :-1].ravel(), g[:, 1:].ravel())[0, 1]) < 4 / np.sqrt(n)
    assert abs(np.corrcoef(g[:-1].ravel(), g[1:].ravel())[0, 1]) < 4 / np.sqrt(n)
