"""GPU parity of the tuple operator (tg_tuple_*: the reference's own e-ph
operator, ScatteringChannel.cpp:421-538,618-688 and IntegratorCore.hpp:260-357)
against the oracle's restatement (oracle/mc_oracle.c).

Bar:
  - the tuple list is identical (same tuples, same order): both sides make the
    same inside/outside decision per MC point;
  - one evaluation agrees to 1e-12 of max|out| (the GPU sums precomputed
    per-tuple moments, the oracle sums the MC points per stage);
  - particle number sum_e int out_e = 0 to 1e-13 of sum |int out_e|;
  - a member of a batch is bitwise equal to its single evaluation.
"""
import os
import struct

import numpy as np
import pytest

import oracle
import tortoise_amd as tg
from tortoise_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _case(dim, ns, seed=0, nmc=50, nu=1):
    w = wl.Workload(f"t{dim}d{ns}", dim, (ns,) * dim)
    eps = w.eps()
    pts = oracle.mc_points(dim, 3, nmc, seed)
    return w, eps, pts


def _state(w, batch, seed):
    rng = np.random.default_rng(seed)
    base = oracle.mc_nodal_to_dg(w.dim, w.nsec, w.initial_f())
    F = np.stack([base + 0.02 * rng.standard_normal(base.shape) for _ in range(batch)])
    bath = rng.uniform(0.05, 0.2, batch)
    return F, bath


def _oracle_out(w, eps, pts, idx, F, n, nu=1):
    fph = np.zeros_like(F)
    fph[:, 0] = n
    return oracle.mc_apply_eph(w.dim, w.nsec, w.origin, w.extent, eps, w.omega0, w.g ** 2, nu, pts,
                               idx, F, fph, nthreads=8)


def _number(w, out):
    vol = abs(np.prod(np.array(w.extent[:w.dim]) / np.array(w.nsec[:w.dim]))) / (2 if w.dim == 2 else 6)
    M = np.linalg.inv(oracle.mc_inverse_mass(w.dim, vol))
    return out @ M[:, 0]       # int of each element's linear form


@pytest.mark.parametrize("dim,ns", [(2, 8), (2, 16), (3, 3)])
def test_tuple_list_and_propagate_match_oracle(dim, ns):
    print(f"[gpu] tuple operator {dim}D ns={ns}", flush=True)
    w, eps, pts = _case(dim, ns)
    idx_o = oracle.mc_tuples_eph(dim, w.nsec, w.origin, w.extent, eps, w.omega0, 1, pts)
    op = tg.TupleOperator.create(w, max_batch=3)
    op.set_points(pts)
    idx_g = op.indices()
    assert idx_g.shape == idx_o.shape and np.array_equal(idx_g, idx_o)
    F, bath = _state(w, 3, seed=ns)
    out = op.propagate(F, bath)
    for b in range(3):
        ref = _oracle_out(w, eps, pts, idx_o, F[b], bath[b])
        assert np.abs(out[b] - ref).max() <= 1e-12 * np.abs(ref).max(), b
        num = _number(w, out[b])
        assert abs(num.sum()) <= 1e-13 * np.abs(num).sum()
    # members of a batch equal single evaluations
    for b in (0, 2):
        assert np.array_equal(op.propagate(F[b], bath[b]), out[b])
    op.close()


HEX_SIDES = np.array([[1.0, 0.0], [-0.5, np.sqrt(3.0) / 2.0]])   # examples/main.cpp:37-39


@pytest.mark.parametrize("ns,nu", [(12, 1), (16, 2)])
def test_hexagonal_zone_matches_oracle(ns, nu):
    """The hexagonal zone of the reference's propagation example (a Region
    from the origin with sides (1, 0) and (-1/2, sqrt(3)/2), meshed whole,
    umklapp vectors over its sides up to the second neighbouring zones as the
    example's channels use): the GPU's tuple list is the oracle's and one
    evaluation agrees to 1e-12 of max|out|; number balances to 1e-13."""
    print(f"[gpu] tuple operator hexagonal zone ns={ns} nu={nu}", flush=True)
    w = wl.Workload(f"hex{ns}", 2, (ns, ns))
    i0, i1 = np.meshgrid(np.arange(ns + 1), np.arange(ns + 1), indexing="xy")
    X = np.outer(i0.ravel() / ns, HEX_SIDES[0]) + np.outer(i1.ravel() / ns, HEX_SIDES[1])
    c = 0.5 * (HEX_SIDES[0] + HEX_SIDES[1])
    eps = 1.2 * ((X - c) ** 2).sum(1)
    pts = oracle.mc_points(2, 3, 50, 7)
    idx_o = oracle.mc_tuples_eph(2, w.nsec, (0.0, 0.0), None, eps, w.omega0, nu, pts, sides=HEX_SIDES)
    op = tg.TupleOperator.create(w, n_umklapp=nu, max_batch=2, eps=eps, origin=(0.0, 0.0), sides=HEX_SIDES)
    op.set_points(pts)
    idx_g = op.indices()
    assert idx_g.shape == idx_o.shape and np.array_equal(idx_g, idx_o)
    rng = np.random.default_rng(ns)
    f = 0.5 + 0.4 * np.tanh((0.4 - eps) / 0.1)
    base = oracle.mc_nodal_to_dg(2, w.nsec, f)
    F = np.stack([base + 0.01 * rng.standard_normal(base.shape) for _ in range(2)])
    bath = np.array([0.08, 0.15])
    out = op.propagate(F, bath)
    vol = abs(np.linalg.det(HEX_SIDES)) / (ns * ns) / 2.0
    M = np.linalg.inv(oracle.mc_inverse_mass(2, vol))
    for b in range(2):
        fph = np.zeros_like(F[b])
        fph[:, 0] = bath[b]
        ref = oracle.mc_apply_eph(2, w.nsec, (0.0, 0.0), None, eps, w.omega0, w.g ** 2, nu, pts, idx_o,
                                  F[b], fph, nthreads=8, sides=HEX_SIDES)
        assert np.abs(out[b] - ref).max() <= 1e-12 * np.abs(ref).max(), b
        num = out[b] @ M[:, 0]
        assert abs(num.sum()) <= 1e-13 * np.abs(num).sum()
    op.close()


def test_degenerate_sides_and_zone_refused():
    """Degenerate mesh sides are refused; so is a degenerate zone when umklapp
    vectors are asked for (they would repeat tuples), not without them."""
    w = wl.Workload("hex4", 2, (4, 4))
    eps = np.zeros(25)
    flat = np.array([[1.0, 0.0], [2.0, 0.0]])
    with pytest.raises(RuntimeError, match="mesh sides are degenerate"):
        tg.TupleOperator.create(w, n_umklapp=0, eps=eps, origin=(0.0, 0.0), sides=flat)
    with pytest.raises(RuntimeError, match="zone sides are degenerate"):
        tg.TupleOperator.create(w, n_umklapp=1, eps=eps, origin=(0.0, 0.0), sides=HEX_SIDES, zone=flat)
    tg.TupleOperator.create(w, n_umklapp=0, eps=eps, origin=(0.0, 0.0), sides=HEX_SIDES, zone=flat).close()


def test_tuple_config0_square32():
    """The reference's own CPU-runnable case: 32 x 32 sections (2048
    triangles), NMC = 50, umklapp {-1,0,1}^2."""
    print("[gpu] tuple operator square32", flush=True)
    w, eps, pts = _case(2, 32)
    idx_o = oracle.mc_tuples_eph(2, w.nsec, w.origin, w.extent, eps, w.omega0, 1, pts)
    op = tg.TupleOperator.create(w, max_batch=1)
    op.set_points(pts)
    assert np.array_equal(op.indices(), idx_o)
    F, bath = _state(w, 1, seed=7)
    out = op.propagate(F[0], bath[0])
    ref = _oracle_out(w, eps, pts, idx_o, F[0], bath[0])
    assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max()
    op.close()


def test_redraw_keeps_tuples_and_matches_oracle():
    """MaterialTimeStatus::propagate redraws the MC points: the tuple list
    stays, the moments follow the new points."""
    w, eps, pts = _case(2, 12, seed=0)
    idx = oracle.mc_tuples_eph(2, w.nsec, w.origin, w.extent, eps, w.omega0, 1, pts)
    op = tg.TupleOperator.create(w, max_batch=1)
    op.set_points(pts)
    F, bath = _state(w, 1, seed=3)
    out0 = op.propagate(F[0], bath[0])
    pts1 = oracle.mc_points(2, 3, 50, seed=99)
    op.set_points(pts1)
    assert np.array_equal(op.indices(), idx)
    out1 = op.propagate(F[0], bath[0])
    ref1 = _oracle_out(w, eps, pts1, idx, F[0], bath[0])
    assert np.abs(out1 - ref1).max() <= 1e-12 * np.abs(ref1).max()
    assert np.abs(out1 - out0).max() > 1e-6 * np.abs(out0).max()
    op.close()


def test_scatt_indices_file_round_trip(tmp_path):
    """.scattIndices layout (ScatteringChannel.cpp:422-459,520-537): per leg
    and axis an int section count, an unsigned long tuple count, then per
    tuple three int element ids and an int umklapp id."""
    w, eps, pts = _case(2, 6)
    op = tg.TupleOperator.create(w, max_batch=1)
    op.set_points(pts)
    idx = op.indices()
    path = str(tmp_path / "[(el:C)(pn:opt)]-[(el:C)].scattIndices")
    op.save_indices(path)
    raw = open(path, "rb").read()
    head = struct.unpack("<6i", raw[:24])
    assert head == (6, 6, 6, 6, 6, 6)
    (count,) = struct.unpack("<Q", raw[24:32])
    assert count == len(idx)
    body = np.frombuffer(raw[32:], np.int32).reshape(-1, 4)
    assert np.array_equal(body, idx)
    # a second operator loads the list instead of enumerating, same result
    op2 = tg.TupleOperator.create(w, max_batch=1)
    op2.load_indices(path)
    op2.set_points(pts)
    F, bath = _state(w, 1, seed=4)
    assert np.array_equal(op2.propagate(F[0], bath[0]), op.propagate(F[0], bath[0]))
    # a file for another mesh is refused
    w8 = wl.Workload("t8", 2, (8, 8))
    op3 = tg.TupleOperator.create(w8, max_batch=1)
    with pytest.raises(tg.TgError, match="sections differ"):
        op3.load_indices(path)
    for o in (op, op2, op3):
        o.close()


def test_set_indices_rejects_bad_lists():
    w, eps, pts = _case(2, 4)
    op = tg.TupleOperator.create(w, max_batch=1)
    nel = 32
    with pytest.raises(tg.TgError, match="element"):
        op.set_indices(np.array([[0, 1, nel, 0]], np.int32))
    with pytest.raises(tg.TgError, match="umklapp"):
        op.set_indices(np.array([[0, 1, 2, 9]], np.int32))
    # a list out of leg-0 order is accepted and kept in leg-0 order
    op.set_indices(np.array([[3, 1, 2, 0], [2, 1, 2, 0]], np.int32))
    assert np.array_equal(op.indices(), [[2, 1, 2, 0], [3, 1, 2, 0]])
    with pytest.raises(tg.TgError, match="MC points"):
        op.propagate(np.zeros((nel, 3)), 0.1)
    op.close()


def test_tuple_dp5_matches_oracle_and_batches_bitwise():
    """MaterialTimeStatus::propagate with the reference operator: DP5(4) of
    the DG state on the device against the oracle's DP5 with the same tuple
    operator (identical step sequence, dense outputs to 1e-10); a member of a
    batch follows exactly its single run (bitwise)."""
    print("[gpu] tuple dp5 8x8 t=5", flush=True)
    w, eps, pts = _case(2, 8)
    idx = oracle.mc_tuples_eph(2, w.nsec, w.origin, w.extent, eps, w.omega0, 1, pts)
    op = tg.TupleOperator.create(w, max_batch=3)
    op.set_points(pts)
    F, bath = _state(w, 3, seed=5)
    fg, dg, st = op.dp5(F[0], 0.0, 5.0, 0.5, 1e-8, 1e-12, bath[0], max_dense=10)
    r = oracle.dp5_mc_eph(2, w.nsec, w.origin, w.extent, eps, w.omega0, w.g ** 2, 1, pts, idx,
                          bath[0], F[0], 0.0, 5.0, 0.5, 1e-8, 1e-12, max_dense=10)
    assert st["attempted"] == r["stats"].attempted and st["accepted"] == r["stats"].accepted
    assert st["n_dense"] == r["stats"].n_dense == len(dg)
    for a, b in zip(dg, r["dense"]):
        assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()
    num = _number(w, fg).sum()
    assert abs(num - _number(w, F[0]).sum()) <= 1e-12 * abs(num)
    fb, db, sb = op.dp5(F, 0.0, 5.0, 0.5, 1e-8, 1e-12, bath, max_dense=10)
    assert np.array_equal(fb[0], fg) and sb[0]["attempted"] == st["attempted"]
    for b in (1, 2):
        f1, _, s1 = op.dp5(F[b], 0.0, 5.0, 0.5, 1e-8, 1e-12, bath[b])
        assert np.array_equal(fb[b], f1) and sb[b]["attempted"] == s1["attempted"]
    op.close()


@pytest.mark.parametrize("dim,ns", [(2, 12), (3, 3)])
def test_scattering_rates_match_oracle(dim, ns):
    """ScatteringChannel::scatteringRates of the electron band (legs 0 and 2)
    from the same per-tuple moments, against the oracle's restatement."""
    print(f"[gpu] tuple rates {dim}D ns={ns}", flush=True)
    w, eps, pts = _case(dim, ns)
    idx = oracle.mc_tuples_eph(dim, w.nsec, w.origin, w.extent, eps, w.omega0, 1, pts)
    op = tg.TupleOperator.create(w, max_batch=2)
    op.set_points(pts)
    F, bath = _state(w, 2, seed=11)
    R = op.rates(F, bath)
    for b in range(2):
        fph = np.zeros_like(F[b])
        fph[:, 0] = bath[b]
        ref = oracle.mc_rates_eph(dim, w.nsec, w.origin, w.extent, eps, w.omega0, w.g ** 2, 1, pts,
                                  idx, F[b], fph, nthreads=8)
        assert np.abs(R[b] - ref).max() <= 1e-12 * np.abs(ref).max(), b
        assert np.array_equal(op.rates(F[b], bath[b]), R[b])
    op.close()


def _acoustic(w):
    X = oracle.mesh_nodes(w.dim, list(w.nsec), list(w.origin), list(w.extent))
    return 0.05 + 0.02 * np.sqrt((X ** 2).sum(1))


def test_dispersive_phonon_bath_matches_oracle():
    """A dispersive phonon band (omega = 0.05 + 0.02 |k|) held as a bath: the
    phonon energy forms enter D, N and the energy windows."""
    print("[gpu] tuple operator, dispersive phonons", flush=True)
    w, eps, pts = _case(2, 12)
    om = _acoustic(w)
    idx = oracle.mc_tuples_eph(2, w.nsec, w.origin, w.extent, eps, w.omega0, 1, pts, eps_ph=om)
    op = tg.TupleOperator.create(w, max_batch=2, eps_ph=om)
    op.set_points(pts)
    assert np.array_equal(op.indices(), idx)
    F, bath = _state(w, 2, seed=21)
    out = op.propagate(F, bath)
    for b in range(2):
        fph = np.zeros_like(F[b])
        fph[:, 0] = bath[b]
        ref = oracle.mc_apply_eph(2, w.nsec, w.origin, w.extent, eps, w.omega0, w.g ** 2, 1, pts, idx,
                                  F[b], fph, nthreads=8, eps_ph=om)
        assert np.abs(out[b] - ref).max() <= 1e-12 * np.abs(ref).max(), b
    op.close()


@pytest.mark.parametrize("dim,ns", [(2, 10), (3, 3)])
def test_dynamical_phonons_match_oracle(dim, ns):
    """The phonon band evolves: both bands' derivatives against the oracle,
    batch members bitwise equal to single runs, and DP5 of both bands with
    the max-over-bands error norm (identical step sequence)."""
    print(f"[gpu] tuple operator, dynamical phonons {dim}D ns={ns}", flush=True)
    w, eps, pts = _case(dim, ns)
    om = _acoustic(w)
    idx = oracle.mc_tuples_eph(dim, w.nsec, w.origin, w.extent, eps, w.omega0, 1, pts, eps_ph=om)
    op = tg.TupleOperator.create(w, max_batch=3, eps_ph=om, dynamic=True)
    op.set_points(pts)
    assert np.array_equal(op.indices(), idx)
    rng = np.random.default_rng(ns)
    fe = oracle.mc_nodal_to_dg(dim, w.nsec, w.initial_f())
    fp = oracle.mc_nodal_to_dg(dim, w.nsec, 1.0 / np.expm1(om / 0.05))
    S = np.stack([np.stack([fe + 0.01 * rng.standard_normal(fe.shape),
                            fp * (1 + 0.1 * rng.standard_normal(fp.shape))]) for _ in range(3)])
    out = op.propagate(S)
    for b in range(3):
        el, ph = oracle.mc_apply_eph(dim, w.nsec, w.origin, w.extent, eps, w.omega0, w.g ** 2, 1, pts,
                                     idx, S[b, 0], S[b, 1], nthreads=8, phonon_out=True, eps_ph=om)
        assert np.abs(out[b, 0] - el).max() <= 1e-12 * np.abs(el).max(), b
        assert np.abs(out[b, 1] - ph).max() <= 1e-12 * np.abs(ph).max(), b
    for b in (0, 2):
        assert np.array_equal(op.propagate(S[b]), out[b])
    if dim == 2:
        fg, dg, st = op.dp5(S[0], 0.0, 5.0, 0.5, 1e-8, 1e-12, max_dense=10)
        r = oracle.dp5_mc_eph(2, w.nsec, w.origin, w.extent, eps, w.omega0, w.g ** 2, 1, pts, idx, 0.0,
                              S[0].reshape(-1, dim + 1), 0.0, 5.0, 0.5, 1e-8, 1e-12, max_dense=10,
                              dynamic=True, eps_ph=om)
        assert st["attempted"] == r["stats"].attempted and st["accepted"] == r["stats"].accepted
        for a, bref in zip(dg, r["dense"]):
            assert np.abs(a.reshape(-1, dim + 1) - bref).max() <= 1e-10 * np.abs(bref).max()
        fb, _, sb = op.dp5(S[:2], 0.0, 5.0, 0.5, 1e-8, 1e-12)
        assert np.array_equal(fb[0], fg) and sb[0]["attempted"] == st["attempted"]
    op.close()


def test_scattering_rates_with_dynamical_phonons_match_oracle():
    w, eps, pts = _case(2, 10)
    om = _acoustic(w)
    idx = oracle.mc_tuples_eph(2, w.nsec, w.origin, w.extent, eps, w.omega0, 1, pts, eps_ph=om)
    op = tg.TupleOperator.create(w, max_batch=2, eps_ph=om, dynamic=True)
    op.set_points(pts)
    fe = oracle.mc_nodal_to_dg(2, w.nsec, w.initial_f())
    fp = oracle.mc_nodal_to_dg(2, w.nsec, 1.0 / np.expm1(om / 0.05))
    R = op.rates(np.stack([fe, fp]))
    ref = oracle.mc_rates_eph(2, w.nsec, w.origin, w.extent, eps, w.omega0, w.g ** 2, 1, pts, idx,
                              fe, fp, nthreads=8, eps_ph=om)
    assert np.abs(R - ref).max() <= 1e-12 * np.abs(ref).max()
    op.close()
