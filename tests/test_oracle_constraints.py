"""Pins of the oracle's constrained bath, step vetoer and bath energy ledger.

The reference enforces a constrained band's population at every stage time
(MaterialStatus::applyConstraints, src/Physics/Status/MaterialStatus.cpp:198-206,
called through enforceValue in DormandPrince54.hpp:548,565-578) and caps the
step with MaterialTimeStatus::timeStepVetoer (MaterialTimeStatus.cpp:60-100,
applied at DormandPrince54.hpp:562-564,640-642).  The checks below use cases
whose answer follows from those rules by hand (a pulse between samples, a
step inside the 1.3 |L - R| allowance, a deviation just below and above
atol + rtol |average|), an independent interpolation (numpy.interp), an
independent integrator (scipy's DOP853) and the exact energy identity the
ledger must satisfy.
"""
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

import oracle

PI = math.pi


def _pulse(t, t_p, width, amp, base=0.1):
    return base + amp * np.exp(-0.5 * ((t - t_p) / width) ** 2)


def _sched(fn, t_end=20.0, dt=1e-3, **kw):
    ts = np.arange(0.0, t_end + dt / 2, dt)
    return oracle.BathSchedule(fn(ts), 0.0, dt, **kw)


# ------------------------------------------------------------ schedule
def test_schedule_is_linear_interpolation_held_at_the_ends():
    rng = np.random.default_rng(3)
    s = rng.random(11)
    sc = oracle.BathSchedule(s, 2.0, 0.5)
    knots = 2.0 + 0.5 * np.arange(11)
    ts = np.concatenate([rng.uniform(0.0, 9.0, 200), knots])
    got = np.array([sc.value(t) for t in ts])
    assert np.abs(got - np.interp(ts, knots, s)).max() <= 1e-15
    assert sc.value(-5.0) == s[0] and sc.value(100.0) == s[-1]


# -------------------------------------------------------------- vetoer
def test_vetoer_without_constraint_returns_t_plus_h_minus_t():
    rng = np.random.default_rng(0)
    for t, h in zip(rng.uniform(0, 100, 50), rng.uniform(1e-3, 5, 50)):
        assert oracle.step_vetoer(None, t, h, 0.5) == (t + h) - t
    assert oracle.step_vetoer(None, 0.0, 0.7, 0.5) == 0.7


def test_vetoer_keeps_steps_over_constant_and_linear_populations():
    # a constant band never varies; a linear one deviates from an interval's
    # average by at most |L - R| / 2 < 1.3 |L - R|
    const = _sched(lambda t: 0.3 + 0.0 * t)
    lin = _sched(lambda t: 0.1 + 0.05 * t)
    for sc in (const, lin):
        for t, h in ((0.0, 4.0), (1.3, 2.5), (7.0, 0.9)):
            assert oracle.step_vetoer(sc, t, h, 0.25) == (t + h) - t


def test_vetoer_stops_at_the_first_sample_on_a_pulse():
    # t = 0, h = 4: timesToTest = 0, 0.8, 1.2, 3.2, 3.56, 4; samples at
    # 0.25, 0.5, ...  A pulse at 1.6 (width 0.05) is 0.068 above the baseline
    # at the sample 1.5 and invisible at 1.25 and at the interval ends.
    sc = _sched(lambda t: _pulse(t, 1.6, 0.05, 0.5))
    assert oracle.step_vetoer(sc, 0.0, 4.0, 0.25) == 1.5
    # the same pulse seen from t = 1.0: samples at 1.25, 1.5, ... -> 0.5
    assert oracle.step_vetoer(sc, 1.0, 4.0, 0.25) == 0.5
    # a pulse between two samples (1.625, width 0.01) is not seen at all
    sc2 = _sched(lambda t: _pulse(t, 1.625, 0.01, 0.5))
    assert oracle.step_vetoer(sc2, 0.0, 4.0, 0.25) == 4.0
    # a step shorter than the dense spacing has no samples
    assert oracle.step_vetoer(sc, 1.4, 0.2, 0.25) == (1.4 + 0.2) - 1.4


def test_vetoer_allows_1p3_times_the_end_difference():
    # a smooth rise 0.1 -> 0.6 inside [1.2, 3.2]: |v - avg| <= 0.25 stays
    # below 1.3 * |L - R| = 0.65, so the step stands
    sc = _sched(lambda t: 0.1 + 0.25 * (1.0 + np.tanh((t - 2.2) / 0.3)))
    assert oracle.step_vetoer(sc, 0.0, 4.0, 0.25) == 4.0


def test_vetoer_threshold_is_atol_plus_rtol_times_average():
    # flat ends (L = R = 0.1): the allowance is atol + rtol * 0.1 = 1.1e-3
    # for atol = rtol = 1e-3; a bump of 1.0e-3 at the sample 1.5 passes,
    # one of 1.2e-3 stops the step there
    for amp, want in ((1.0e-3, 4.0), (1.2e-3, 1.5)):
        sc = _sched(lambda t: _pulse(t, 1.5, 0.02, amp), atol=1e-3, rtol=1e-3)
        assert oracle.step_vetoer(sc, 0.0, 4.0, 0.25) == want
    # a looser absolute tolerance lets the larger bump pass too
    sc = _sched(lambda t: _pulse(t, 1.5, 0.02, 1.2e-3), atol=2e-3, rtol=1e-3)
    assert oracle.step_vetoer(sc, 0.0, 4.0, 0.25) == 4.0


# --------------------------------------------------- DP5 with a schedule
@pytest.fixture(scope="module")
def k8():
    nsec = [8, 8]
    origin, ext = [-PI, -PI], [2 * PI, 2 * PI]
    X = oracle.mesh_nodes(2, nsec, origin, ext)
    eps = 0.5 * (X ** 2).sum(1)
    kern = oracle.Kernel(2, nsec, origin, ext, eps, g=0.1, omega0=0.05, sigma=0.01)
    f0 = 1 / (np.exp((eps - 0.5) / 0.05) + 1) + 0.2 * np.exp(-0.5 * ((eps - 0.9) / 0.03) ** 2)
    return kern, kern.full(), eps, f0


def test_dp5_constant_schedule_equals_constant_bath(k8):
    kern, K, eps, f0 = k8
    a = oracle.dp5_collision(2, kern.nsec, K, kern.V, 0.0894, f0, 0.0, 10.0, 0.5, 1e-8, 1e-12,
                             max_dense=20)
    sc = _sched(lambda t: 0.0894 + 0.0 * t)
    b = oracle.dp5_collision(2, kern.nsec, K, kern.V, 0.5, f0, 0.0, 10.0, 0.5, 1e-8, 1e-12,
                             max_dense=20, sched=sc)
    assert np.array_equal(a["h"], b["h"]) and np.array_equal(a["dense"], b["dense"])


def test_dp5_pulsed_bath_matches_an_independent_integrator(k8):
    """f(t) under a bath pulse n(t) equals scipy's DOP853 on the same rhs, and
    the vetoer stopped a step at the pulse."""
    kern, K, eps, f0 = k8
    pulse = lambda t: _pulse(t, 3.0, 0.3, 2.0, base=0.0894)  # noqa: E731
    sc = _sched(pulse, t_end=10.0, dt=1e-3)
    r = oracle.dp5_collision(2, kern.nsec, K, kern.V, 0.0, f0, 0.0, 10.0, 0.5, 1e-9, 1e-13,
                             max_dense=20, sched=sc)
    assert r["stats"].n_dense == 20

    def rhs(t, y):
        return oracle.collision_apply(K, kern.V, sc.value(t), y)

    ref = solve_ivp(rhs, (0.0, 10.0), f0, method="DOP853", rtol=1e-12, atol=1e-14,
                    t_eval=r["dense_t"])
    err = np.abs(r["dense"] - ref.y.T).max() / np.abs(f0).max()
    assert err < 1e-7, err
    # the bath pulse changes the trajectory (the check is not vacuous)
    c = oracle.dp5_collision(2, kern.nsec, K, kern.V, 0.0894, f0, 0.0, 10.0, 0.5, 1e-9, 1e-13,
                             max_dense=20)
    assert np.abs(c["dense"][-1] - r["dense"][-1]).max() > 1e-4
    # resolving the pulse takes more accepted steps than the constant bath
    assert r["accepted"].sum() > c["accepted"].sum()


def test_ledger_closes_the_energy_balance(k8):
    """E_el(t_end) + ledger = E_el(0) to roundoff.  y5 = y + sum_j (h b_j) k_j,
    so the identity holds exactly only when the ledger sums the flux of the
    right stages with the step's b weights and carries k7's flux into the next
    step's k1 (FSAL); a bh weight, a dropped stage or a stale k1 breaks it."""
    kern, K, eps, f0 = k8
    r = oracle.dp5_collision(2, kern.nsec, K, kern.V, 0.0894, f0, 0.0, 10.0, 0.05, 1e-10, 1e-14,
                             max_dense=200, eps=eps)
    e0 = np.dot(kern.V * eps, f0)
    e1 = np.dot(kern.V * eps, r["y"])
    assert abs(e1 + r["ledger"] - e0) <= 1e-13 * e0
    assert abs(r["ledger"]) > 1e-3 * e0   # energy did move


def test_ledger_with_a_pulsed_bath(k8):
    # the balance holds with a time-dependent bath and number-defect removal
    kern, K, eps, f0 = k8
    sc = _sched(lambda t: _pulse(t, 3.0, 0.3, 2.0, base=0.0894), t_end=10.0)
    r = oracle.dp5_collision(2, kern.nsec, K, kern.V, 0.0, f0, 0.0, 10.0, 0.5, 1e-8, 1e-12,
                             max_dense=20, sched=sc, eps=eps, conserve=True)
    e0 = np.dot(kern.V * eps, f0)
    assert abs(np.dot(kern.V * eps, r["y"]) + r["ledger"] - e0) <= 1e-13 * e0


# ----------------------------------------------- tuple operator (DG states)
@pytest.fixture(scope="module")
def mc6():
    nsec = [6, 6]
    origin, ext = [-PI, -PI], [2 * PI, 2 * PI]
    X = oracle.mesh_nodes(2, nsec, origin, ext)
    eps = 0.5 * (X ** 2).sum(1)
    pts = oracle.mc_points(2, 3, 50, 1)
    idx = oracle.mc_tuples_eph(2, nsec, origin, ext, eps, 0.05, 1, pts)
    fn = 1 / (np.exp((eps - 1.0) / 0.5) + 1)
    return nsec, origin, ext, eps, pts, idx, oracle.mc_nodal_to_dg(2, nsec, fn)


def test_energy_dg_closed_form(mc6):
    nsec, origin, ext, eps, pts, idx, fdg = mc6
    # int eps * 1 over the box = int |k|^2/2 of the P1 interpolant of eps;
    # for f = 1 this is sum_i V_i eps_i (lumped = consistent for a constant)
    one = oracle.mc_nodal_to_dg(2, nsec, np.ones_like(eps))
    V = oracle.node_volumes(2, nsec, ext)
    assert abs(oracle.energy_dg(2, nsec, ext, eps, one) - np.dot(V, eps)) <= 1e-13 * np.dot(V, eps)


def test_tuple_dp5_constant_schedule_and_ledger(mc6):
    nsec, origin, ext, eps, pts, idx, fdg = mc6
    args = (2, nsec, origin, ext, eps, 0.05, 0.01, 1, pts, idx)
    a = oracle.dp5_mc_eph(*args, 0.0894, fdg, 0.0, 4.0, 0.2, 1e-8, 1e-12, max_dense=20)
    sc = _sched(lambda t: 0.0894 + 0.0 * t, t_end=4.0)
    b = oracle.dp5_mc_eph(*args, 0.7, fdg, 0.0, 4.0, 0.2, 1e-8, 1e-12, max_dense=20, sched=sc)
    assert np.array_equal(a["dense"], b["dense"]) and a["ledger"] == b["ledger"]
    e0 = oracle.energy_dg(2, nsec, ext, eps, fdg)
    e1 = oracle.energy_dg(2, nsec, ext, eps, a["f"])
    assert abs(e1 + a["ledger"] - e0) <= 1e-13 * abs(e0)
    assert abs(a["ledger"]) > 1e-4 * abs(e0)


def test_replay_reproduces_the_ledger(k8):
    # the fixed-step replay of a run's accepted steps gives its ledger and
    # final state bit for bit (same stages, same flux sums, same order)
    kern, K, eps, f0 = k8
    sc = _sched(lambda t: _pulse(t, 3.0, 0.3, 2.0, base=0.0894), t_end=10.0)
    r = oracle.dp5_collision(2, kern.nsec, K, kern.V, 0.0, f0, 0.0, 10.0, 0.5, 1e-8, 1e-10,
                             max_dense=20, sched=sc, eps=eps)
    rp = oracle.dp5_collision_replay(2, kern.nsec, K, kern.V, 0.0, f0, 0.0, 10.0, 0.5, r["h"],
                                     r["accepted"], sched=sc, eps=eps)
    assert rp["ledger"] == r["ledger"] and np.array_equal(rp["y"], r["y"])
