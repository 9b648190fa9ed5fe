"""GPU parity of the hat-averaged occupation factors (TG_OCC_HAT) against the
CPU oracle.

With TG_OCC_HAT the W-form's occupation factor of the test node is the hat
average fbar_i = int phi_i f / int phi_i instead of the nodal f_i; the
products with K still use the nodal f.  Checks: one FP64 evaluation (2D and
3D, with and without the number-defect removal), the batched mixed-precision
evaluation against FP64, the LDS-row and global-memory epilogues bit for bit,
the adaptive FP64 DP5 against the oracle's (identical steps), the batched
mixed DP5 against FP64 replays with number and energy (electrons + bath
ledger) conservation, and the adjoint (with and without the number-defect
removal in its forward map) against the oracle adjoint.
Tolerances as in the nodal tests: 1e-13 (FP64), 2e-6 (mixed evaluation),
1e-12 (DP5 dense outputs), 1e-6 (mixed DP5 vs replay), 1e-10 (adjoint).
"""
import math

import numpy as np
import pytest
import torch  # before the library: one HIP runtime per process

import oracle
import tortoise_amd as tg
from tortoise_amd import autograd as tga
from tortoise_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _members(w, B, seed=0):
    eps = w.eps()
    centre, amp, temp = wl.ensemble_members(B, seed)
    F = np.stack([wl.initial_distribution(eps, w.mu, temp[b], centre[b], w.pump_width, amp[b])
                  for b in range(B)], axis=1)
    bath = np.array([wl.bose_einstein(w.omega0, t) for t in temp])
    return F, bath


@pytest.fixture(scope="module")
def sq32_hat():
    w = wl.SQUARE32
    print("[gpu] hat occupation square32 B=8", flush=True)
    ctx = tg.Context.create(w, max_batch=8, occupation="hat")
    ctx.assemble(mixed=True)
    ok = oracle.Kernel(w.dim, w.nsec, w.origin, w.extent, w.eps(), w.g, w.omega0, w.sigma)
    return w, ctx, ok, ok.full(), oracle.HatStencil(2, w.nsec)


def test_hat_apply_matches_oracle(sq32_hat):
    w, ctx, ok, K, S = sq32_hat
    F, bath = _members(w, 3, seed=11)
    for conserve in (False, True):
        Ig = ctx.apply(F, bath_n=bath, conserve=conserve)
        for b in range(3):
            Io = oracle.collision_apply(K, ok.V, bath[b], F[:, b], conserve=conserve, hat=S)
            assert _rel(Ig[:, b], Io) <= 1e-13, (conserve, b, _rel(Ig[:, b], Io))
        if conserve:
            assert np.abs(ok.V @ Ig).max() <= 1e-14 * np.abs(ok.V[:, None] * Ig).sum(0).max()
    # the nodal context differs (the factors matter)
    cn = tg.Context.create(w, max_batch=1)
    cn.assemble()
    assert _rel(cn.apply(F[:, 0], bath_n=bath[:1]), Ig[:, 0]) > 1e-6
    cn.close()


def test_hat_apply_cube_matches_oracle():
    w = wl.Workload("cube6", 3, (6, 6, 6), sigma=0.02)
    print(f"[gpu] hat occupation cube6 n={w.n_nodes}", flush=True)
    ctx = tg.Context.create(w, max_batch=1, occupation="hat")
    ctx.assemble()
    ok = oracle.Kernel(3, w.nsec, w.origin, w.extent, w.eps(), w.g, w.omega0, w.sigma)
    S = oracle.HatStencil(3, w.nsec)
    f0 = w.initial_f()
    for conserve in (False, True):
        Io = oracle.collision_apply(ok.full(), ok.V, w.bath_n, f0, conserve=conserve, hat=S)
        assert _rel(ctx.apply(f0, conserve=conserve), Io) <= 1e-13


def test_hat_mixed_matches_fp64(sq32_hat):
    w, ctx, ok, K, S = sq32_hat
    F, bath = _members(w, 8)
    for conserve in (False, True):
        Im = ctx.apply(F, bath_n=bath, mixed=True, conserve=conserve)
        Ip = ctx.apply(F, bath_n=bath, conserve=conserve)
        for b in range(8):
            assert _rel(Im[:, b], Ip[:, b]) <= 2e-6, (conserve, b, _rel(Im[:, b], Ip[:, b]))


def test_hat_mixed_rows_outside_lds(monkeypatch):
    """The global-memory pack and epilogue compute the same occupation
    factors, bit for bit, as the LDS-row ones."""
    w = wl.SQUARE32
    F, bath = _members(w, 8, seed=7)
    for k, v in [("TG_MIXED_NW", "64"), ("TG_MIXED_WAVES", "4"), ("TG_MIXED_FOLD", "1"),
                 ("TG_MIXED_DEPTH", "2"), ("TG_MIXED_SPLITS", "1")]:
        monkeypatch.setenv(k, v)
    ca = tg.Context.create(w, max_batch=8, occupation="hat")
    ca.assemble(mixed=True)
    Ia = ca.apply(F, bath_n=bath, mixed=True, conserve=True)
    monkeypatch.setenv("TG_MIXED_ROWCACHE", "0")
    cb = tg.Context.create(w, max_batch=8, occupation="hat")
    cb.assemble(mixed=True)
    Ib = cb.apply(F, bath_n=bath, mixed=True, conserve=True)
    assert np.array_equal(Ia, Ib)


def test_hat_dp5_matches_oracle(sq32_hat):
    w, ctx, ok, K, S = sq32_hat
    f0 = w.initial_f()
    print("[gpu] hat dp5 square32 t=100", flush=True)
    fg, dg, sg = ctx.dp5(f0, 0.0, w.t_final, w.dense_dt, w.rtol, w.atol, conserve=True, max_dense=100)
    ro = oracle.dp5_collision(2, w.nsec, K, ok.V, w.bath_n, f0, 0.0, w.t_final, w.dense_dt,
                              w.rtol, w.atol, conserve=True, max_dense=100, occupation="hat")
    so = ro["stats"]
    assert (sg["attempted"], sg["accepted"], sg["rejected"]) == (so.attempted, so.accepted, so.rejected)
    assert max(_rel(dg[k], ro["dense"][k]) for k in range(100)) <= 1e-12
    n0, _ = ctx.conservation(f0)
    n1, _ = ctx.conservation(dg[-1])
    assert abs(n1 - n0) <= 1e-12 * n0


def test_hat_mixed_dp5_vs_fp64_replay(sq32_hat):
    """Batched mixed DP5 (the fused member passes) with hat factors against
    FP64 replays of each member's accepted steps; number and energy
    (electrons + bath ledger) conservation."""
    w, ctx, ok, K, S = sq32_hat
    F, bath = _members(w, 8, seed=3)
    print("[gpu] hat mixed dp5 square32 B=8 t=100", flush=True)
    fm, dm, stats, steps = ctx.dp5(F, 0.0, 100.0, 1.0, 1e-6, 1e-12, bath_n=bath, mixed=True,
                                   conserve=True, max_dense=100, h_cap=4096)
    assert all(s["n_dense"] == 100 and not s["failed"] for s in stats)
    fr, dr = ctx.dp5_replay(F, 0.0, 100.0, 1.0, steps, bath_n=bath, conserve=True, max_dense=100)
    eps = w.eps()
    for b in range(8):
        assert _rel(dm[b, -1], dr[b, -1]) <= 1e-6, (b, _rel(dm[b, -1], dr[b, -1]))
        n0 = ok.V @ F[:, b]
        assert abs(ok.V @ dm[b, -1] - n0) <= 1e-12 * n0
        e0 = ok.V @ (eps * F[:, b])
        e1 = ok.V @ (eps * fm[:, b]) + stats[b]["bath_energy"]
        assert abs(e1 - e0) <= 1e-12 * e0, (b, (e1 - e0) / e0)
    # member 0's FP64 replay against the oracle's replay of the same steps
    ro = oracle.dp5_collision_replay(2, w.nsec, K, ok.V, bath[0], F[:, 0], 0.0, 100.0, 1.0, steps[0],
                                     np.ones(len(steps[0]), bool), conserve=True, occupation="hat")
    assert _rel(dr[0, -1], ro["dense"][-1]) <= 1e-12


@pytest.mark.parametrize("occ,conserve", [("hat", False), ("hat", True), ("nodal", True)])
def test_adjoint_matches_oracle_adjoint(occ, conserve):
    """The GPU adjoint with hat factors and/or the number-defect removal in
    every stage against the oracle's adjoint of the same fixed-step map."""
    w = wl.Workload("square32_adj", 2, (32, 32), g=0.12, omega0=0.045, t_final=4.0)
    print(f"[gpu] adjoint vs oracle, square32 t=4, {occ} factors, conserve={conserve}", flush=True)
    T = 0.02
    bath, dbath = tga.bath_occupation(w.omega0, T), tga.dbath_domega0(w.omega0, T)
    ctx = tg.Context.create(w, max_batch=1, bath_n=bath, occupation=occ)
    ctx.assemble()
    f0 = w.initial_f()
    _, _, _, steps = ctx.dp5(f0, 0.0, 4.0, 0.04, 1e-8, 1e-12, bath_n=bath, max_dense=100, h_cap=4096,
                             conserve=conserve)
    dL = np.random.default_rng(0).normal(size=ctx.n)
    g_f0, g_g, g_w = ctx.dp5_adjoint(f0, 0.0, 4.0, 0.04, steps, dL, bath, dbath, ckpt=5, conserve=conserve)
    ok = oracle.Kernel(2, w.nsec, w.origin, w.extent, w.eps(), w.g, w.omega0, w.sigma)
    o_f0, o_g, o_w = oracle.dp5_adjoint(ok.full(), ok.domega_full(), ok.V, bath, dbath, w.g, f0,
                                        0.0, 4.0, 0.04, steps, dL, ckpt=7,
                                        hat=oracle.HatStencil(2, w.nsec) if occ == "hat" else None,
                                        conserve=conserve)
    assert np.linalg.norm(g_f0 - o_f0) <= 1e-10 * np.linalg.norm(o_f0)
    assert abs(g_g - o_g) <= 1e-10 * abs(o_g) and abs(g_w - o_w) <= 1e-10 * abs(o_w)
