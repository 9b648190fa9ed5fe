"""GPU parity of the electron-electron (4-leg) channel against the oracle's
restatement (oracle/ee_oracle.c), on 2D and 3D lattices and on the full
20^3 cube (8000 periodic points, sigma_ee = 0.02).

Tolerances: the GPU accumulates the pair spectra as int64 fixed point (exact
and order-independent) and converts once, so one evaluation agrees with the
FP64 oracle to ~1e-12 relative; tests allow 1e-10.  Number conservation
sum_i V_i I_i = 0 holds to FP64 roundoff.
"""
import numpy as np
import pytest

import oracle
import tortoise_amd as tg
from tortoise_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _members(w, B, seed):
    eps = w.eps()
    centre, amp, temp = wl.ensemble_members(B, seed)
    F = np.stack([wl.initial_distribution(eps, w.mu, temp[b], centre[b], w.pump_width, amp[b])
                  for b in range(B)], axis=1)
    bath = np.array([wl.bose_einstein(w.omega0, t) for t in temp])
    return F, bath


def _ee_pair(w, B):
    """Contexts with and without the e-e channel; the difference of their
    outputs is the e-e operator."""
    with_ee = tg.Context.create(w, max_batch=B)
    with_ee.assemble()
    without = tg.Context.create(w, max_batch=B, ee_g=0.0)
    without.assemble()
    return with_ee, without


@pytest.mark.parametrize("dim,ns", [(2, 12), (3, 6)])
def test_ee_matches_oracle_small(dim, ns):
    w = wl.Workload(f"ee{dim}d", dim, (ns,) * dim, ee_g=2.0, ee_sigma=0.3, sigma=0.3)
    print(f"[gpu] e-e {dim}D ns={ns}", flush=True)
    F, bath = _members(w, 3, seed=dim)
    ce, cp = _ee_pair(w, 3)
    Iee = ce.apply(F, bath_n=bath) - cp.apply(F, bath_n=bath)
    for b in range(3):
        Io = oracle.ee_apply(dim, w.nsec, w.extent, w.eps(), F[:, b], 2.0, 0.09, 0.3, 0.3 / 4)
        total = ce.apply(F[:, [b]], bath_n=bath[[b]])[:, 0]
        assert np.abs(Iee[:, b] - Io).max() <= 1e-10 * np.abs(total).max(), b
        # the e-e part is a sizeable share of the operator in this setting
        assert np.linalg.norm(Io) > 1e-2 * np.linalg.norm(total)


def test_ee_batch_members_equal_single_evaluations():
    w = wl.Workload("ee3d", 3, (6, 6, 6), ee_g=2.0, ee_sigma=0.3, sigma=0.3)
    F, bath = _members(w, 4, seed=9)
    ce = tg.Context.create(w, max_batch=4)
    ce.assemble()
    Ib = ce.apply(F, bath_n=bath)
    for b in (0, 3):
        assert np.array_equal(ce.apply(F[:, [b]], bath_n=bath[[b]])[:, 0], Ib[:, b])


def test_ee_cube20_matches_oracle_and_conserves_number():
    w = wl.CUBE20
    print("[gpu] e-e cube20 (8000 points, sigma 0.02)", flush=True)
    F, bath = _members(w, 2, seed=1)
    ce, cp = _ee_pair(w, 2)
    V = ce.node_volumes()
    Iee = ce.apply(F, bath_n=bath) - cp.apply(F, bath_n=bath)
    Io = oracle.ee_apply(3, w.nsec, w.extent, w.eps(), F[:, 0], w.ee_g, w.ee_kappa2, w.ee_sigma,
                         w.ee_sigma / 4)
    assert np.abs(Iee[:, 0] - Io).max() <= 1e-10 * np.abs(Io).max()
    # the e-e operator alone conserves particle number exactly
    assert abs(V @ Io) <= 1e-12 * (V @ np.abs(Io))
    for b in range(2):
        assert abs(V @ Iee[:, b]) <= 1e-10 * (V @ np.abs(Iee[:, b]))


def test_ee_cube20_batched_dp5():
    """e-ph (mixed) plus e-e through the batched DP5 (attempt graph with the
    e-e kernels): every member finishes, number drift <= 1e-12, and the
    members agree with an FP64 replay of their own steps."""
    w = wl.CUBE20
    B = 4
    F, bath = _members(w, B, seed=2)
    ce = tg.Context.create(w, max_batch=B)
    ce.assemble(mixed=True)
    V = ce.node_volumes()
    print("[gpu] e-e cube20 dp5 B=4 t=2", flush=True)
    fm, _, stats, steps = ce.dp5(F, 0.0, 2.0, 0.02, 1e-6, 1e-12, bath_n=bath, mixed=True,
                                 conserve=True, h_cap=4096)
    assert all(not s["failed"] for s in stats)
    fr, _ = ce.dp5_replay(F, 0.0, 2.0, 0.02, steps, bath_n=bath)
    for b in range(B):
        n0 = V @ F[:, b]
        assert abs(V @ fm[:, b] - n0) <= 1e-12 * n0
        assert np.linalg.norm(fm[:, b] - fr[:, b]) <= 1e-6 * np.linalg.norm(fr[:, b])


def test_adjoint_refuses_ee_channel():
    w = wl.Workload("ee2d", 2, (8, 8), ee_g=1.0, ee_sigma=0.3, sigma=0.3)
    ce = tg.Context.create(w, max_batch=1)
    ce.assemble()
    f0 = w.initial_f()
    with pytest.raises(tg.TgError, match="e-e"):
        ce.dp5_adjoint(f0, 0.0, 1.0, 0.01, np.full(10, 0.1), np.ones(ce.n), 0.09, 0.0, ckpt=4)
