"""The oracle's 4-leg channel {el in, el in, el out, el out} (mc_oracle.c:
or_mc_tuples_ee, or_mc_tuple_ee, or_mc_apply_ee; the reference's
ScatteringChannel<NDim,4>, ScatteringChannel.cpp:285-415,421-538,618-670 and the
NLEGS == 4 branch of IntegratorCore.hpp:260-357), pinned by what the
mathematics fixes rather than by retyping its formulas:

  - pruning: the box- and energy-pruned list equals the brute-force list
    (every leg-3 element tested);
  - the delta shell: on every sampled tuple sum_l dir_l int val eps_l(K_l) = 0
    and sum_l dir_l int val k_l = -G u int val, with eps_l and k_l built here
    from the mesh's nodes, independent of the oracle's D and N (a wrong sign,
    index or column in D, N or the constrained solve fails them);
  - the MC weight: with val = 1 the integral equals an independent numpy
    estimate of int dK0 dK1 dK2' [inside] / |det D_C| (Dirichlet-sampled
    simplex points, numpy's solve) within 5 standard errors;
  - the forms: an empty band (f = 0) and a full band (f = 1) give exactly zero
    (Pauli blocking of both products);
  - particle number: sum_e int out_e = 0 for any state.
"""
import numpy as np
import pytest

import oracle
from tortoise_amd import workloads as wl


def _mesh(dim, ns):
    w = wl.Workload(f"ee4_{dim}d{ns}", dim, (ns,) * dim)
    nn = ns + 1
    axes = [np.linspace(w.origin[d], w.origin[d] + w.extent[d], nn) for d in range(dim)]
    grids = np.meshgrid(*axes, indexing="ij")
    # node index ix + iy (n+1) (+ iz (n+1)^2)
    nodes = np.stack([g.transpose(*range(dim - 1, -1, -1)).ravel() for g in grids], 1)
    conn = np.zeros((int(np.prod(w.nsec[:dim])) * (2 if dim == 2 else 6), dim + 1), np.int32)
    oracle.lib().or_mesh_elements(dim, np.ascontiguousarray(w.nsec, np.int32), conn)
    return w, nodes, conn


def _forms(w, nodes, conn, eps, e):
    v = conn[e]
    ef = np.concatenate([[eps[v[0]]], eps[v[1:]] - eps[v[0]]])
    kf = np.stack([np.concatenate([[nodes[v[0], d]], nodes[v[1:], d] - nodes[v[0], d]]) for d in range(w.dim)])
    return ef, kf


def _umklapp(dim, uid, nu=1):
    span = 2 * nu + 1
    u = []
    for _ in range(dim):
        u.append(uid % span - nu)
        uid //= span
    return np.array(u[::-1], float)


def test_pruned_list_equals_brute_force():
    w = wl.Workload("ee4_2d3", 2, (3, 3))
    pts = oracle.mc_points(2, 4, 50, 11)
    a = oracle.mc_tuples_ee(2, w.nsec, w.origin, w.extent, w.eps(), 1, pts)
    b = oracle.mc_tuples_ee(2, w.nsec, w.origin, w.extent, w.eps(), 1, pts, brute=True)
    assert a.shape[0] > 1000 and np.array_equal(a, b)
    # (e0, e1, e2, uid, e3) order
    key = a[:, [0, 1, 2, 4, 3]].astype(np.int64)
    assert all(tuple(key[i]) < tuple(key[i + 1]) for i in range(len(key) - 1))


@pytest.mark.parametrize("dim,ns", [(2, 4), (3, 2)])
def test_delta_shell_conserves_energy_and_momentum(dim, ns):
    w, nodes, conn = _mesh(dim, ns)
    eps = w.eps()
    pts = oracle.mc_points(dim, 4, 50, 5)
    nu = 1 if dim == 2 else 0
    idx = oracle.mc_tuples_ee(dim, w.nsec, w.origin, w.extent, eps, nu, pts)
    assert idx.shape[0] > 100
    rng = np.random.default_rng(dim)
    dirs = [1, 1, -1, -1]
    nc = dim + 1
    checked = 0
    for tup in idx[rng.choice(idx.shape[0], 200, replace=False)]:
        c = rng.uniform(0.1, 0.9, (4, nc)) * np.r_[1.0, [0.1] * dim]
        res, _, _, ok = oracle.mc_tuple_ee(dim, w.nsec, w.origin, w.extent, eps, nu, pts, tup, c)
        assert ok
        sE, sK, scale = 0.0, np.zeros(dim), 0.0
        for l in range(4):
            ef, kf = _forms(w, nodes, conn, eps, tup[l])
            sE += dirs[l] * ef @ res[l]
            sK += dirs[l] * kf @ res[l]
            scale += np.abs(ef) @ np.abs(res[l]) + np.abs(kf).sum(0) @ np.abs(res[l])
            assert res[l][0] == res[0][0]        # every leg receives the same basis-0 integral
        G = np.array(w.extent[:dim]) * _umklapp(dim, tup[4], nu)
        assert abs(sE) <= 1e-13 * scale
        assert np.abs(sK + G * res[0][0]).max() <= 1e-13 * scale
        checked += res[0][0] != 0.0
    assert checked > 50


def _simplex(rng, d, n):
    """Uniform points in the reference d-simplex (Dirichlet(1,..,1) minus the
    last coordinate): an independent sampler."""
    if d == 1:
        return rng.uniform(0.0, 1.0, (n, 1))
    x = rng.exponential(size=(n, d + 1))
    return (x / x.sum(1, keepdims=True))[:, :d]


def test_mc_weight_matches_independent_estimate():
    dim, ns = 2, 4
    w, nodes, conn = _mesh(dim, ns)
    eps = w.eps()
    nmc = 20000
    pts = oracle.mc_points(dim, 4, nmc, 9)
    idx = oracle.mc_tuples_ee(dim, w.nsec, w.origin, w.extent, eps, 1, oracle.mc_points(dim, 4, 50, 9))
    rng = np.random.default_rng(4)
    # val = 1: f = 0 on the in-legs, 1 on the out-legs
    c = np.zeros((4, 3))
    c[2, 0] = c[3, 0] = 1.0
    dirs = np.array([1.0, 1.0, -1.0, -1.0])
    done = 0
    for tup in idx[rng.choice(idx.shape[0], 60, replace=False)]:
        res, _, _, ok = oracle.mc_tuple_ee(dim, w.nsec, w.origin, w.extent, eps, 1, pts, tup, c)
        if not ok or res[0][0] == 0.0:
            continue
        # independent: D, N from the nodes; K0, K1 in the triangle, K2's first
        # coordinate in [0, 1]; solve D_C [K2_1, K3] = N - D_F K_F
        D = np.zeros((3, 8))
        N = np.zeros(3)
        for l in range(4):
            ef, kf = _forms(w, nodes, conn, eps, tup[l])
            D[:2, 2 * l:2 * l + 2] = dirs[l] * kf[:, 1:]
            D[2, 2 * l:2 * l + 2] = dirs[l] * ef[1:]
            N[:2] -= dirs[l] * kf[:, 0]
            N[2] -= dirs[l] * ef[0]
        N[:2] -= np.array(w.extent[:2]) * _umklapp(2, tup[4])
        n = 400000
        K0, K1 = _simplex(rng, 2, n), _simplex(rng, 2, n)
        k2a = rng.uniform(0.0, 1.0, n)
        rhs = N[None, :] - K0 @ D[:, 0:2].T - K1 @ D[:, 2:4].T - np.outer(k2a, D[:, 4])
        sol = np.linalg.solve(D[:, 5:8], rhs.T).T
        K2 = np.stack([k2a, sol[:, 0]], 1)
        K3 = sol[:, 1:]
        inside = ((K2 > 0).all(1) & (K2.sum(1) < 1) & (K3 > 0).all(1) & (K3.sum(1) < 1)).astype(float)
        detc = abs(np.linalg.det(D[:, 5:8]))
        # int over K0, K1 (area 1/2 each) and K2's free coordinate (length 1)
        est = inside.mean() * 0.25 / detc
        se = inside.std() / np.sqrt(n) * 0.25 / detc
        se_o = np.sqrt(res[0][0] * 0.25 / detc / nmc)    # the oracle's own MC error (hits ~ Poisson)
        assert abs(res[0][0] - est) <= 5 * np.hypot(se, se_o) + 1e-15, (tup, res[0][0], est)
        done += 1
        if done == 8:
            break
    assert done >= 4


@pytest.mark.parametrize("kappa", [0.0, 0.6])
def test_apply_pauli_blocking_and_number(kappa):
    dim, ns = 2, 4
    w = wl.Workload("ee4_2d4", dim, (ns, ns))
    eps = w.eps()
    pts = oracle.mc_points(dim, 4, 50, 2)
    idx = oracle.mc_tuples_ee(dim, w.nsec, w.origin, w.extent, eps, 1, pts)
    nel = 2 * ns * ns
    F = np.zeros((4, nel, 3))
    F[1, :, 0] = 1.0                                  # full band
    rng = np.random.default_rng(1)
    F[2] = rng.uniform(0.05, 0.6, (nel, 3)) * np.array([1.0, 0.2, 0.2])
    F[3] = rng.uniform(0.3, 0.7, (nel, 3)) * np.array([1.0, 0.1, 0.1])
    out = oracle.mc_apply_ee(dim, w.nsec, w.origin, w.extent, eps, 1, pts, idx, F, amplitude=1.3, kappa=kappa)
    assert not np.any(out[0]) and not np.any(out[1])
    vol = (w.extent[0] / ns) * (w.extent[1] / ns) / 2
    for b in (2, 3):
        num = vol * (out[b][:, 0] + (out[b][:, 1] + out[b][:, 2]) / 3)
        assert np.abs(out[b]).max() > 1e-6
        assert abs(num.sum()) <= 1e-13 * np.abs(num).sum()
