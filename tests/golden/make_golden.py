"""Writes the golden fixtures of tests/golden/ from the CPU oracle.

    python tests/golden/make_golden.py

Fixtures (numpy .npz, no pickles):
  square32_kernel.npz   rows 0, 300, 544, 1088 of K; I(f0); V; the 1089-node
                        32x32-section square at g=0.1, omega0=0.05, sigma=0.01
  square32_dp5.npz      first 10 accepted steps and the dense outputs at
                        t = 1..10 of the adaptive DP5 (rtol 1e-8, atol 1e-12)
  ensemble256.npz       the 256 member draws (pump centre, amplitude, T; seed 0)
  square8_gradients.npz central finite-difference gradients (h = 1e-6
                        relative) of the energy-binned loss wrt g and omega0
                        on the 8x8-section square, with the frozen steps
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from tortoise_amd import workloads as wl  # noqa: E402


def square32():
    w = wl.SQUARE32
    eps = w.eps()
    k = oracle.Kernel(2, w.nsec, w.origin, w.extent, eps, w.g, w.omega0, w.sigma)
    K = k.full()
    f0 = w.initial_f()
    I = oracle.collision_apply(K, k.V, w.bath_n, f0)
    rows = np.array([0, 300, 544, 1088])
    np.savez(os.path.join(HERE, "square32_kernel.npz"), rows=rows, K_rows=K[rows], I=I, V=k.V,
             f0=f0)
    r = oracle.dp5_collision(2, w.nsec, K, k.V, w.bath_n, f0, 0.0, 10.0, 1.0, w.rtol, w.atol,
                             max_dense=10)
    acc = r["h"][r["accepted"]]
    np.savez(os.path.join(HERE, "square32_dp5.npz"), h=acc[:10], dense=r["dense"],
             dense_t=r["dense_t"], attempted=r["stats"].attempted)


def ensemble():
    c, a, t = wl.ensemble_members(256, seed=0)
    np.savez(os.path.join(HERE, "ensemble256.npz"), centre=c, amplitude=a, temperature=t)


def gradients():
    T = 0.02
    nsec = [8, 8]
    pi = math.pi
    X = oracle.mesh_nodes(2, nsec, [-pi, -pi], [2 * pi, 2 * pi])
    eps = 0.5 * (X ** 2).sum(1)
    g0, w0 = 0.12, 0.045

    def run(g, om, steps=None):
        k = oracle.Kernel(2, nsec, [-pi, -pi], [2 * pi, 2 * pi], eps, g, om, 0.01)
        f0 = 1 / (np.exp((eps - 0.5) / 0.02) + 1) + 0.2 * np.exp(-0.5 * ((eps - 0.9) / 0.03) ** 2)
        nb = 1.0 / math.expm1(om / T)
        if steps is None:
            r = oracle.dp5_collision(2, nsec, k.full(), k.V, nb, f0, 0.0, 4.0, 0.4, 1e-8, 1e-12)
            return r["h"][r["accepted"]], k.V
        r = oracle.dp5_collision_replay(2, nsec, k.full(), k.V, nb, f0, 0.0, 4.0, 0.4, steps,
                                        np.ones(len(steps), bool))
        return r["dense"][-1]

    steps, V = run(g0, w0)
    edges = np.linspace(0.0, 2.0, 13)
    Wb = np.stack([np.where((eps >= a) & (eps < b), V, 0.0) for a, b in zip(edges[:-1], edges[1:])])
    f_ref = run(g0, w0, steps)
    target = Wb @ (0.97 * f_ref)

    def loss(g, om):
        d = Wb @ run(g, om, steps) - target
        return float(d @ d)

    hg, hw = 1e-6 * g0, 1e-6 * w0
    dg = (loss(g0 + hg, w0) - loss(g0 - hg, w0)) / (2 * hg)
    dw = (loss(g0, w0 + hw) - loss(g0, w0 - hw)) / (2 * hw)
    np.savez(os.path.join(HERE, "square8_gradients.npz"), steps=steps, target=target, dL_dg=dg,
             dL_domega0=dw, g=g0, omega0=w0, edges=edges)


if __name__ == "__main__":
    square32()
    ensemble()
    gradients()
    print("fixtures written to", HERE)
