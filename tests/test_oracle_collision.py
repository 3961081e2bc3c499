"""Single-cell pins of the oracle's collision and LES maths (closed forms and
invariants; none of them restates the oracle's formulas)."""
import numpy as np
import pytest

import oracle


def axes():
    d, opp, w = oracle.tables()
    return np.stack([d[0], d[2], d[1]], axis=1), opp, w  # e[q] = (x, y, z)


def random_state(rng, amp=0.02):
    rho = 0.95 + 0.1 * rng.random()
    u = 0.08 * (rng.random(3) - 0.5)
    f = oracle.cell_feq(rho, u) * (1.0 + amp * rng.standard_normal(19))
    e, _, _ = axes()
    r = f.sum()
    return f, r, (f @ e) / r


def test_mrt_equals_srt_when_all_rates_are_one_over_tau():
    """With every rate 1/tau and RefDen = rho, M^-1 S (M f - meq) = (f - feq)/tau."""
    rng = np.random.default_rng(2)
    for _ in range(50):
        f, rho, u = random_state(rng)
        tau = 0.5 + rng.random()
        te = 0.1 * rng.random()
        p = oracle.default_params(tau=tau, ref_den=rho, s_rates=[-1.0] * 19)
        a = oracle.cell_collide(p, f, rho, u, te, mrt=True)
        b = oracle.cell_collide(p, f, rho, u, te, mrt=False)
        assert np.abs(a - b).max() < 1e-15


@pytest.mark.parametrize("mrt", [True, False])
def test_collision_conserves_mass_and_momentum(mrt):
    e, _, _ = axes()
    rng = np.random.default_rng(3)
    for _ in range(50):
        f, rho, u = random_state(rng, amp=0.05)
        p = oracle.default_params(tau=0.6, smagorinsky=0.17)
        te = oracle.cell_tau_eddy(p, f, rho, u)
        out = oracle.cell_collide(p, f, rho, u, te, mrt=mrt)
        assert abs(out.sum() - rho) < 1e-14
        assert np.allclose(out @ e, rho * u, atol=1e-15)


@pytest.mark.parametrize("mrt", [True, False])
def test_rest_state_is_a_fixed_point(mrt):
    _, _, w = axes()
    p = oracle.default_params(tau=0.6, smagorinsky=0.17)
    f = w.copy()
    assert oracle.cell_tau_eddy(p, f, 1.0, np.zeros(3)) == 0.0
    out = oracle.cell_collide(p, f, 1.0, np.zeros(3), 0.0, mrt=mrt)
    assert np.abs(out - w).max() < 1e-16


def test_each_moment_relaxes_at_its_own_rate():
    """Perturb one non-conserved moment k: M f* = meq + (1 - s_k)(m - meq) e_k.
    Pins the pairing of rows of M with columns of M^-1 and every rate of S."""
    m = oracle.mrt_matrix()
    _, minv = oracle.mrt_inverse_exact(m)
    rng = np.random.default_rng(4)
    rho, u, tau, te = 1.03, np.array([0.02, -0.015, 0.01]), 0.62, 0.03
    rates = dict(zip(range(19), [0, 1.19, 1.4, 0, 1.2, 0, 1.2, 0, 1.2, None, 1.4, None, 1.4, None, None, None,
                                 1.98, 1.98, 1.98]))
    feq = oracle.cell_feq(rho, u)
    p = oracle.default_params(tau=tau, ref_den=rho)  # RefDen = rho makes meq = M feq
    for k in range(19):
        if k in (0, 3, 5, 7):
            continue  # conserved moments
        c = 1e-3 * (1 + rng.random())
        f = feq + minv[:, k] * c
        out = oracle.cell_collide(p, f, rho, u, te, mrt=True)
        s_k = 1.0 / (tau + te) if rates[k] is None else rates[k]
        expect = feq + (1.0 - s_k) * minv[:, k] * c
        assert np.abs(out - expect).max() < 1e-15, k


def test_les_vanishes_for_cs_zero():
    rng = np.random.default_rng(5)
    f, rho, u = random_state(rng, amp=0.1)
    p = oracle.default_params(tau=0.6, smagorinsky=0.0)
    assert abs(oracle.cell_tau_eddy(p, f, rho, u)) < 1e-16


def test_les_stress_and_relaxation_time_closed_form():
    """f = feq + w_q (9/2)(e_a e_b - d_ab/3) A_ab has non-equilibrium stress
    Pi = A exactly (D3Q19 isotropy); the total relaxation time t = tau + tau_eddy
    is then the positive root of t^2 - tau t - (18 sqrt2/4) Cs^2 |A| / rho = 0."""
    e, _, w = axes()
    rng = np.random.default_rng(6)
    for _ in range(20):
        rho = 0.97 + 0.06 * rng.random()
        u = 0.06 * (rng.random(3) - 0.5)
        a = 1e-3 * rng.standard_normal((3, 3))
        a = a + a.T
        h = (4.5 * w) * (np.einsum("qa,qb,ab->q", e, e, a) - np.trace(a) / 3.0)
        f = oracle.cell_feq(rho, u) + h
        tau, cs = 0.55 + 0.2 * rng.random(), 0.1 + 0.1 * rng.random()
        p = oracle.default_params(tau=tau, smagorinsky=cs)
        te = oracle.cell_tau_eddy(p, f, rho, u)
        t = tau + te
        qnorm = np.sqrt((a * a).sum())
        assert abs(t * t - tau * t - 18.0 * 1.41421356237 / 4.0 * cs * cs * qnorm / rho) < 1e-14
        assert te > 0.0
