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


@pytest.mark.parametrize("dim,ns", [(2, 8), (3, 5)])
def test_adjoint_with_ee_channel_against_central_fd(dim, ns):
    """dL/df0 and dL/dg through the fixed-step DP5 with both channels (e-ph
    W-form plus e-e) against central finite differences of the same fixed-step
    map (tg_dp5_replay), for L = dL . f(t1).  The e-e vector-Jacobian product
    (tg_ee.hip, MODES 1 and 2, with the sum of mu over each lattice point's
    2, 4 or 8 boundary nodes) is pinned by these differences alone: the
    oracle's adjoint covers the e-ph channel."""
    from tortoise_amd import autograd as tga
    T = 0.02
    w = wl.Workload(f"ee{dim}d_adj", dim, (ns,) * dim, ee_g=1.5, ee_sigma=0.3, sigma=0.3)
    bath, dbath = tga.bath_occupation(w.omega0, T), tga.dbath_domega0(w.omega0, T)
    t1, dt = 2.0, 0.05
    nd = int(round(t1 / dt))

    def ctx_at(g, ee_g=w.ee_g):
        c = tg.Context.create(w, max_batch=1, g=g, bath_n=bath, ee_g=ee_g)
        c.assemble()
        return c

    ctx = ctx_at(w.g)
    f0 = w.initial_f()
    print(f"[gpu] adjoint with e-e, {w.name} t=2", flush=True)
    _, _, st, steps = ctx.dp5(f0, 0.0, t1, dt, 1e-8, 1e-12, bath_n=bath, max_dense=nd, h_cap=4096)
    assert st["accepted"] > 5
    rng = np.random.default_rng(4)
    dL = rng.normal(size=ctx.n)
    g_f0, g_g, _ = ctx.dp5_adjoint(f0, 0.0, t1, dt, steps, dL, bath, dbath, ckpt=5)

    def L_of(c, f):
        _, dense = c.dp5_replay(f, 0.0, t1, dt, steps, bath_n=bath, max_dense=nd)
        return float(dL @ dense[-1])

    # the e-e channel moves the loss by far more than the tolerance
    assert abs(L_of(ctx, f0) - L_of(ctx_at(w.g, ee_g=0.0), f0)) > 1e-2 * abs(L_of(ctx, f0))
    for k in range(3):
        v = rng.normal(size=ctx.n)
        e = 1e-6
        fd = (L_of(ctx, f0 + e * v) - L_of(ctx, f0 - e * v)) / (2 * e)
        assert abs(g_f0 @ v - fd) <= 1e-6 * abs(fd), (k, g_f0 @ v, fd)
    hg = 1e-6 * w.g
    fd_g = (L_of(ctx_at(w.g + hg), f0) - L_of(ctx_at(w.g - hg), f0)) / (2 * hg)
    assert abs(g_g - fd_g) <= 1e-6 * abs(fd_g), (g_g, fd_g)
