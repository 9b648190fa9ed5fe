"""GPU parity of the 4-leg tuple operator (tg_tuple4_create: the reference's
ScatteringChannel<NDim,4> on the electron band, ScatteringChannel.cpp:285-415,
421-538,618-670 and IntegratorCore.hpp:260-357) against the oracle's
restatement (oracle/mc_oracle.c: or_mc_tuples_ee, or_mc_apply_ee).

Bar:
  - the tuple list is identical (same tuples, same order): both sides make the
    same inside/outside decision per MC point;
  - one evaluation agrees to 1e-12 of max|out| (both sum the MC points per
    tuple and each element's entries in tuple order);
  - particle number sum_e int out_e = 0 to 1e-13 of sum |int out_e|;
  - a member of a batch is bitwise equal to its single evaluation;
  - a list in another order (one sweep per umklapp vector, as the reference
    writes it) and a .scattIndices round trip give the oracle's result on
    that list;
  - DP5 with the 4-leg operator: the same attempted and accepted steps as the
    oracle's DP5, dense outputs to 1e-10, particle number to 1e-12, members
    of a batch bitwise equal to single runs.
"""
import numpy as np
import pytest

import oracle
import tortoise_amd as tg
from tortoise_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _states(w, batch, seed):
    rng = np.random.default_rng(seed)
    nel = int(np.prod(w.nsec[:w.dim])) * (2 if w.dim == 2 else 6)
    scale = np.r_[1.0, [0.15] * w.dim]
    return rng.uniform(0.05, 0.7, (batch, nel, w.dim + 1)) * scale


def _number(w, out):
    vol = abs(np.prod(np.array(w.extent[:w.dim]) / np.array(w.nsec[:w.dim]))) / (2 if w.dim == 2 else 6)
    M = np.linalg.inv(oracle.mc_inverse_mass(w.dim, vol))
    return out @ M[:, 0]


@pytest.mark.parametrize("dim,ns,nu,kappa", [(2, 4, 1, 0.0), (2, 5, 1, 0.7), (3, 2, 0, 0.5)])
def test_tuple4_list_and_propagate_match_oracle(dim, ns, nu, kappa):
    print(f"[gpu] 4-leg tuple operator {dim}D ns={ns} nu={nu} kappa={kappa}", flush=True)
    w = wl.Workload(f"ee4_{dim}d{ns}", dim, (ns,) * dim)
    eps = w.eps()
    pts = oracle.mc_points(dim, 4, 50, 17)
    idx_o = oracle.mc_tuples_ee(dim, w.nsec, w.origin, w.extent, eps, nu, pts)
    op = tg.TupleOperator.create(w, n_umklapp=nu, max_batch=3, amplitude=1.0, legs=4, kappa=kappa)
    assert op.rows == pts.shape[0]
    op.set_points(pts)
    idx_g = op.indices()
    assert idx_g.shape == idx_o.shape and np.array_equal(idx_g, idx_o)
    F = _states(w, 3, seed=ns)
    out = op.propagate(F, time_kernel=True)
    ref = oracle.mc_apply_ee(dim, w.nsec, w.origin, w.extent, eps, nu, pts, idx_o, F, amplitude=1.0, kappa=kappa)
    for b in range(3):
        assert np.abs(out[b] - ref[b]).max() <= 1e-12 * np.abs(ref[b]).max(), b
        num = _number(w, out[b])
        assert abs(num.sum()) <= 1e-13 * np.abs(num).sum()
    print(f"[gpu]   {idx_g.shape[0]} tuples, bitwise equal to the oracle: {np.array_equal(out, ref)}, "
          f"{op.last_kernel_ms():.3f} ms", flush=True)
    for b in (0, 2):
        assert np.array_equal(op.propagate(F[b]), out[b])
    op.close()


def test_tuple4_reference_order_and_file_round_trip(tmp_path):
    print("[gpu] 4-leg list in umklapp-major order, .scattIndices round trip", flush=True)
    dim, ns, nu = 2, 4, 1
    w = wl.Workload("ee4_2d4", dim, (ns, ns))
    eps = w.eps()
    pts = oracle.mc_points(dim, 4, 50, 23)
    idx = oracle.mc_tuples_ee(dim, w.nsec, w.origin, w.extent, eps, nu, pts)
    # the reference's order: one full sweep per umklapp vector
    ref_order = idx[np.argsort(idx[:, 4], kind="stable")]
    op = tg.TupleOperator.create(w, n_umklapp=nu, max_batch=2, amplitude=1.0, legs=4)
    op.set_indices(ref_order)
    op.set_points(pts)
    assert np.array_equal(op.indices(), ref_order)
    F = _states(w, 2, seed=3)
    out = op.propagate(F)
    ref = oracle.mc_apply_ee(dim, w.nsec, w.origin, w.extent, eps, nu, pts, ref_order, F, amplitude=1.0)
    assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max()
    path = str(tmp_path / "ee.scattIndices")
    op.save_indices(path)
    raw = np.fromfile(path, np.int32)
    assert np.array_equal(raw[:4 * dim], [ns] * (4 * dim))       # nSecSplits per leg and axis
    assert int(raw[4 * dim:4 * dim + 2].view(np.uint64)[0]) == ref_order.shape[0]
    assert np.array_equal(raw[4 * dim + 2:].reshape(-1, 5), ref_order)
    op2 = tg.TupleOperator.create(w, n_umklapp=nu, max_batch=2, amplitude=1.0, legs=4)
    op2.load_indices(path)
    op2.set_points(pts)
    assert np.array_equal(op2.propagate(F), out)
    with pytest.raises(tg.TgError):
        op.rates(F)
    op.close()
    op2.close()


def test_tuple4_dp5_matches_oracle_and_batches_bitwise():
    print("[gpu] 4-leg tuple dp5 2D ns=3 t=2", flush=True)
    dim, ns, nu, amp = 2, 3, 1, 0.02
    w = wl.Workload("ee4_2d3", dim, (ns, ns))
    eps = w.eps()
    pts = oracle.mc_points(dim, 4, 50, 29)
    idx = oracle.mc_tuples_ee(dim, w.nsec, w.origin, w.extent, eps, nu, pts)
    op = tg.TupleOperator.create(w, n_umklapp=nu, max_batch=2, amplitude=amp, legs=4)
    op.set_points(pts)
    F = _states(w, 2, seed=5)
    fg, dg, st = op.dp5(F[0], 0.0, 2.0, 0.2, 1e-8, 1e-12, max_dense=10)
    r = oracle.dp5_mc_ee(dim, w.nsec, w.origin, w.extent, eps, nu, pts, idx, F[0], 0.0, 2.0, 0.2, 1e-8, 1e-12,
                         amplitude=amp, max_dense=10)
    print(f"[gpu]   attempted {st['attempted']} accepted {st['accepted']}", flush=True)
    assert st["attempted"] == r["stats"].attempted and st["accepted"] == r["stats"].accepted
    assert st["n_dense"] == r["stats"].n_dense == len(dg)
    for a, b in zip(dg, r["dense"]):
        assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()
    assert np.abs(fg - F[0]).max() > 1e-3
    num0, num1 = _number(w, F[0]).sum(), _number(w, fg).sum()
    assert abs(num1 - num0) <= 1e-12 * abs(num0)
    fb, _, sb = op.dp5(F, 0.0, 2.0, 0.2, 1e-8, 1e-12, max_dense=10)
    assert np.array_equal(fb[0], fg) and sb[0]["attempted"] == st["attempted"]
    f1, _, _ = op.dp5(F[1], 0.0, 2.0, 0.2, 1e-8, 1e-12)
    assert np.array_equal(fb[1], f1)
    op.close()
