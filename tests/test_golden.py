"""The oracle and the workload definitions reproduce the committed golden
fixtures (tests/golden/, written by make_golden.py); the GPU path matches
them too (gpu-marked tests)."""
import math
import os

import numpy as np
import pytest

import oracle
from tortoise_amd import workloads as wl

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def test_oracle_reproduces_square32_fixture():
    d = _load("square32_kernel.npz")
    w = wl.SQUARE32
    k = oracle.Kernel(2, w.nsec, w.origin, w.extent, w.eps(), w.g, w.omega0, w.sigma)
    for r, row in zip(d["rows"], d["K_rows"]):
        assert np.abs(k.row(int(r)) - row).max() <= 1e-14 * np.abs(row).max()
    assert np.array_equal(w.initial_f(), d["f0"])
    I = oracle.collision_apply(k.full(), k.V, w.bath_n, d["f0"])
    assert np.linalg.norm(I - d["I"]) <= 1e-14 * np.linalg.norm(d["I"])


def test_oracle_reproduces_dp5_fixture():
    d = _load("square32_dp5.npz")
    w = wl.SQUARE32
    k = oracle.Kernel(2, w.nsec, w.origin, w.extent, w.eps(), w.g, w.omega0, w.sigma)
    r = oracle.dp5_collision(2, w.nsec, k.full(), k.V, w.bath_n, w.initial_f(), 0.0, 10.0, 1.0,
                             w.rtol, w.atol, max_dense=10)
    # the error norm is a difference of nearly equal sums and the oracle's
    # OpenMP reductions follow the thread count: step sizes agree to ~1e-6,
    # the fixed-time dense outputs to the local error level
    assert r["stats"].attempted == int(d["attempted"])
    assert np.allclose(r["h"][r["accepted"]][:10], d["h"], rtol=1e-5, atol=0)
    assert np.abs(r["dense"] - d["dense"]).max() <= 1e-9


def test_ensemble_draws_fixture():
    d = _load("ensemble256.npz")
    c, a, t = wl.ensemble_members(256, seed=0)
    assert np.array_equal(c, d["centre"]) and np.array_equal(a, d["amplitude"])
    assert np.array_equal(t, d["temperature"])


def test_oracle_adjoint_matches_fd_fixture():
    d = _load("square8_gradients.npz")
    T = 0.02
    nsec = [8, 8]
    pi = math.pi
    X = oracle.mesh_nodes(2, nsec, [-pi, -pi], [2 * pi, 2 * pi])
    eps = 0.5 * (X ** 2).sum(1)
    g0, w0 = float(d["g"]), float(d["omega0"])
    k = oracle.Kernel(2, nsec, [-pi, -pi], [2 * pi, 2 * pi], eps, g0, w0, 0.01)
    f0 = 1 / (np.exp((eps - 0.5) / 0.02) + 1) + 0.2 * np.exp(-0.5 * ((eps - 0.9) / 0.03) ** 2)
    nb = 1.0 / math.expm1(w0 / T)
    e = math.exp(w0 / T)
    dnb = -e / (T * (e - 1.0) ** 2)
    steps = d["steps"]
    K = k.full()
    r = oracle.dp5_collision_replay(2, nsec, K, k.V, nb, f0, 0.0, 4.0, 0.4, steps,
                                    np.ones(len(steps), bool))
    edges = d["edges"]
    Wb = np.stack([np.where((eps >= a) & (eps < b), k.V, 0.0) for a, b in zip(edges[:-1], edges[1:])])
    dL = 2 * Wb.T @ (Wb @ r["dense"][-1] - d["target"])
    _, gg, gw = oracle.dp5_adjoint(K, k.domega_full(), k.V, nb, dnb, g0, f0, 0.0, 4.0, 0.4, steps,
                                   dL)
    assert abs(gg - float(d["dL_dg"])) <= 1e-5 * abs(float(d["dL_dg"]))
    assert abs(gw - float(d["dL_domega0"])) <= 1e-5 * abs(float(d["dL_domega0"]))


@pytest.mark.gpu
def test_gpu_matches_square32_fixture():
    import tortoise_amd as tg

    d = _load("square32_kernel.npz")
    w = wl.SQUARE32
    ctx = tg.Context.create(w, max_batch=1)
    ctx.assemble()
    for r, row in zip(d["rows"], d["K_rows"]):
        assert np.abs(ctx.kernel_rows(int(r), 1)[0] - row).max() <= 1e-12 * np.abs(row).max()
    I = ctx.apply(d["f0"])
    assert np.linalg.norm(I - d["I"]) <= 1e-13 * np.linalg.norm(d["I"])
    dd = _load("square32_dp5.npz")
    _, dense, st = ctx.dp5(d["f0"], 0.0, 10.0, 1.0, w.rtol, w.atol, max_dense=10)
    assert st["attempted"] == int(dd["attempted"])
    assert np.abs(dense - dd["dense"]).max() <= 1e-9
