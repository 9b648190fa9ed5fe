"""Pins of the oracle's discrete adjoint against central finite differences
(FP64, h = 1e-6 relative) on the frozen step sequence, on an 8x8-section
mesh: gradients of an energy-binned L2 loss with respect to g, omega0 (with
the bath n_BE(omega0, T)) and the initial distribution."""
import math

import numpy as np
import pytest

import oracle

PI = math.pi
T_BATH = 0.02


def _setup(g, omega0, nsec=(8, 8)):
    X = oracle.mesh_nodes(2, list(nsec), [-PI, -PI], [2 * PI, 2 * PI])
    eps = 0.5 * (X ** 2).sum(1)
    k = oracle.Kernel(2, list(nsec), [-PI, -PI], [2 * PI, 2 * PI], eps, g, omega0, 0.01)
    return k, eps


def _bath(omega0):
    return 1.0 / math.expm1(omega0 / T_BATH)


def _dbath(omega0):
    e = math.exp(omega0 / T_BATH)
    return -e / (T_BATH * (e - 1.0) ** 2)


def _bins(eps, V, nb=12):
    edges = np.linspace(0.0, 2.0, nb + 1)
    W = np.zeros((nb, eps.size))
    for b in range(nb):
        sel = (eps >= edges[b]) & (eps < edges[b + 1])
        W[b, sel] = V[sel]
    return W


@pytest.fixture(scope="module")
def problem():
    g0, w0 = 0.12, 0.045
    k, eps = _setup(g0, w0)
    K = k.full()
    f0 = 1 / (np.exp((eps - 0.5) / 0.02) + 1) + 0.2 * np.exp(-0.5 * ((eps - 0.9) / 0.03) ** 2)
    t_f, dt = 4.0, 0.4
    run = oracle.dp5_collision(2, k.nsec, K, k.V, _bath(w0), f0, 0.0, t_f, dt, 1e-8, 1e-12,
                               max_dense=100)
    steps = run["h"][run["accepted"]]
    Wb = _bins(eps, k.V)
    target = Wb @ (f0 * 0.97)
    return dict(g=g0, w=w0, k=k, eps=eps, K=K, f0=f0, t_f=t_f, dt=dt, steps=steps, Wb=Wb,
                target=target)


def _final(p, g, w, f0):
    k, _ = _setup(g, w)
    r = oracle.dp5_collision_replay(2, k.nsec, k.full(), k.V, _bath(w), f0, 0.0, p["t_f"],
                                    p["dt"], p["steps"], np.ones(len(p["steps"]), bool))
    return r["dense"][-1]


def _loss(p, f):
    d = p["Wb"] @ f - p["target"]
    return float(d @ d)


def test_vjp_against_fd(problem):
    p = problem
    rng = np.random.default_rng(0)
    f = rng.uniform(0.05, 0.95, p["f0"].size)
    mu = rng.normal(size=f.size)
    dirv = rng.normal(size=f.size)
    h = 1e-6
    Ip = oracle.collision_apply(p["K"], p["k"].V, 0.3, f + h * dirv)
    Im = oracle.collision_apply(p["K"], p["k"].V, 0.3, f - h * dirv)
    fd = mu @ (Ip - Im) / (2 * h)
    ad = oracle.collision_vjp(p["K"], p["k"].V, 0.3, f, mu) @ dirv
    assert abs(fd - ad) <= 1e-7 * abs(ad)


def test_adjoint_against_central_fd(problem):
    p = problem
    fT = _final(p, p["g"], p["w"], p["f0"])
    dL_df = 2 * p["Wb"].T @ (p["Wb"] @ fT - p["target"])
    Kd = p["k"].domega_full()
    for ck in (32, 3):
        gf0, gg, gw = oracle.dp5_adjoint(p["K"], Kd, p["k"].V, _bath(p["w"]), _dbath(p["w"]),
                                         p["g"], p["f0"], 0.0, p["t_f"], p["dt"], p["steps"],
                                         dL_df, ckpt=ck)
        hg = 1e-6 * p["g"]
        fdg = (_loss(p, _final(p, p["g"] + hg, p["w"], p["f0"]))
               - _loss(p, _final(p, p["g"] - hg, p["w"], p["f0"]))) / (2 * hg)
        assert abs(gg - fdg) <= 1e-5 * abs(fdg), (gg, fdg)
        hw = 1e-6 * p["w"]
        fdw = (_loss(p, _final(p, p["g"], p["w"] + hw, p["f0"]))
               - _loss(p, _final(p, p["g"], p["w"] - hw, p["f0"]))) / (2 * hw)
        assert abs(gw - fdw) <= 1e-5 * abs(fdw), (gw, fdw)
        for i in (5, 40, 70):
            e = np.zeros_like(p["f0"])
            e[i] = 1e-6
            fdi = (_loss(p, _final(p, p["g"], p["w"], p["f0"] + e))
                   - _loss(p, _final(p, p["g"], p["w"], p["f0"] - e))) / 2e-6
            assert abs(gf0[i] - fdi) <= 1e-5 * max(abs(fdi), 1e-3 * np.abs(gf0).max()), (i, gf0[i], fdi)
