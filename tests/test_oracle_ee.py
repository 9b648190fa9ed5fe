"""Pins of the oracle's electron-electron (4-leg) operator (oracle/ee_oracle.c).

The operator is new (no reference function computes it in this form), so it
is pinned by the mathematics: the brute-force sum against independent Python
loops, exact particle-number conservation, the empty and full bands as fixed
points, cubic dependence on f, convergence of the projected energy delta to
the unprojected one as the grid is refined, and a Pauli-blocked two-state
case with a closed form.
"""
import itertools
import math

import numpy as np
import pytest

import oracle

TWO_PI = 2.0 * math.pi


def _lattice(dim, ns):
    """Mesh of the zone [-pi, pi]^dim with ns sections per axis: node
    energies, lumped masses, and the periodic point of every node."""
    nsec = [ns] * dim
    extent = [TWO_PI] * dim
    X = oracle.mesh_nodes(dim, nsec, [-math.pi] * dim, extent)
    eps = 0.5 * (X ** 2).sum(1)
    V = oracle.node_volumes(dim, nsec, extent)
    idx = np.array(list(itertools.product(range(ns + 1), repeat=dim)))[:, ::-1]  # x fastest
    per = np.zeros(len(eps), np.int64)
    for d in range(dim):
        per += (idx[:, d] % ns) * ns ** d
    return nsec, extent, eps, V, per


def _periodic_random(per, rng, lo=0.0, hi=1.0):
    vals = rng.uniform(lo, hi, per.max() + 1)
    return vals[per]


def _direct_python(dim, ns, eps, f, per, g, kappa2, sigma, cutoff=8.0):
    """The definition, written independently: explicit lattice vectors,
    minimum-image q, brute-force sum over (k2, k3) with k4 = k1 + k2 - k3."""
    P = ns ** dim
    dk = TWO_PI / ns
    e = np.zeros(P)
    ff = np.zeros(P)
    e[per] = eps
    ff[per] = f
    coords = list(itertools.product(range(ns), repeat=dim))
    index = {c[::-1]: i for i, c in enumerate(coords)}  # first axis fastest

    def vec(i):
        return coords[i][::-1]

    w = dk ** dim
    norm = 1.0 / (sigma * math.sqrt(2 * math.pi))
    out = np.zeros(P)
    for p1 in range(P):
        c1 = vec(p1)
        s = 0.0
        for p3 in range(P):
            c3 = vec(p3)
            q2 = 0.0
            for a, b in zip(c1, c3):
                qd = ((a - b) % ns) * dk
                if qd >= math.pi:
                    qd -= TWO_PI
                q2 += qd * qd
            amp = g * g / (q2 + kappa2) ** 2
            for p2 in range(P):
                c2 = vec(p2)
                c4 = tuple((a + b - c) % ns for a, b, c in zip(c1, c2, c3))
                p4 = index[c4]
                x = e[p1] + e[p2] - e[p3] - e[p4]
                if abs(x) > cutoff * sigma:
                    continue
                d = norm * math.exp(-0.5 * x * x / sigma ** 2)
                s += w * w * amp * d * (ff[p3] * ff[p4] * (1 - ff[p1]) * (1 - ff[p2])
                                        - ff[p1] * ff[p2] * (1 - ff[p3]) * (1 - ff[p4]))
        out[p1] = s
    return out[per]


@pytest.mark.parametrize("dim,ns", [(2, 3), (2, 4), (3, 2)])
def test_direct_sum_matches_independent_loops(dim, ns):
    nsec, extent, eps, V, per = _lattice(dim, ns)
    rng = np.random.default_rng(dim * 10 + ns)
    f = _periodic_random(per, rng)
    Io = oracle.ee_apply_direct(dim, nsec, extent, eps, f, 0.7, 0.09, 0.6)
    Ip = _direct_python(dim, ns, eps, f, per, 0.7, 0.09, 0.6)
    assert np.abs(Io - Ip).max() <= 1e-12 * np.abs(Ip).max()


@pytest.mark.parametrize("dim,ns,dw", [(2, 8, 0.05), (3, 4, 0.1)])
def test_number_conservation_and_duplicates(dim, ns, dw):
    nsec, extent, eps, V, per = _lattice(dim, ns)
    rng = np.random.default_rng(3)
    f = _periodic_random(per, rng, -0.1, 1.2)
    for I in (oracle.ee_apply(dim, nsec, extent, eps, f, 1.0, 0.09, 0.4, dw),
              oracle.ee_apply_direct(dim, nsec, extent, eps, f, 1.0, 0.09, 0.4)):
        assert abs(V @ I) <= 1e-13 * (V @ np.abs(I))
        # duplicate boundary nodes carry the value of their periodic image
        for p in np.unique(per):
            vals = I[per == p]
            assert np.all(vals == vals[0])


def test_empty_and_full_bands_are_fixed_points():
    nsec, extent, eps, V, per = _lattice(2, 6)
    for c in (0.0, 1.0):
        f = np.full(eps.size, c)
        assert np.all(oracle.ee_apply(2, nsec, extent, eps, f, 1.0, 0.09, 0.4, 0.05) == 0.0)


def test_cubic_in_f():
    """I(f) is a cubic polynomial in f (the quartic terms cancel): its fourth
    finite difference along any direction vanishes."""
    nsec, extent, eps, V, per = _lattice(2, 6)
    rng = np.random.default_rng(7)
    f = _periodic_random(per, rng, 0.2, 0.8)
    d = _periodic_random(per, rng, -1, 1)
    h = 0.05
    Ik = [oracle.ee_apply(2, nsec, extent, eps, f + k * h * d, 1.0, 0.09, 0.4, 0.05)
          for k in range(5)]
    fourth = Ik[4] - 4 * Ik[3] + 6 * Ik[2] - 4 * Ik[1] + Ik[0]
    third = Ik[3] - 3 * Ik[2] + 3 * Ik[1] - Ik[0]
    assert np.abs(fourth).max() <= 1e-12 * np.abs(Ik[0]).max()
    assert np.abs(third).max() > 1e-6 * np.abs(Ik[0]).max()


def test_projection_converges_to_direct_gaussian():
    """Hat projection of the delta on the grid m*dw: the error against the
    unprojected Gaussian falls ~4x per halving of dw (second order)."""
    nsec, extent, eps, V, per = _lattice(2, 10)
    rng = np.random.default_rng(11)
    f = _periodic_random(per, rng)
    sigma = 0.3
    Id = oracle.ee_apply_direct(2, nsec, extent, eps, f, 1.0, 0.09, sigma)
    errs = []
    for dw in (sigma / 2, sigma / 4, sigma / 8, sigma / 16):
        Ip = oracle.ee_apply(2, nsec, extent, eps, f, 1.0, 0.09, sigma, dw)
        errs.append(np.linalg.norm(Ip - Id) / np.linalg.norm(Id))
    assert errs[-1] < 5e-3
    for a, b in zip(errs, errs[1:]):
        assert b < 0.4 * a, errs


def test_two_level_pauli_blocking():
    """Closed form: on a lattice where only energy-conserving quadruples with
    equal energies contribute (sigma tiny, all energies equal), the delta is
    the Gaussian peak 1/(sigma sqrt(2 pi)) and I_1 = w^2 G(0) sum A [f3 f4 (1-f1)(1-f2)
    - f1 f2 (1-f3)(1-f4)]; with f uniform = c this is 0, and a single state
    raised above the rest relaxes (I < 0 there, I > 0 elsewhere)."""
    dim, ns = 2, 4
    nsec, extent, eps, V, per = _lattice(dim, ns)
    flat = np.zeros_like(eps)
    f = np.full(eps.size, 0.3)
    I0 = oracle.ee_apply_direct(dim, nsec, extent, flat, f, 1.0, 0.09, 0.1)
    assert np.abs(I0).max() <= 1e-14
    hot = per == per[0]
    f2 = f.copy()
    f2[hot] = 0.6
    I1 = oracle.ee_apply_direct(dim, nsec, extent, flat, f2, 1.0, 0.09, 0.1)
    assert np.all(I1[hot] < 0) and np.all(I1[~hot] > 0)
    # closed form for the hot point: sum over (k2, k3) of A(k1-k3) x Pauli terms
    w = (TWO_PI / ns) ** 2
    G0 = 1.0 / (0.1 * math.sqrt(2 * math.pi))
    P = ns ** dim
    dk = TWO_PI / ns
    expect = 0.0
    for p3 in range(P):
        c3 = (p3 % ns, p3 // ns)
        q2 = sum((((-c) % ns) * dk - (TWO_PI if ((-c) % ns) * dk >= math.pi else 0.0)) ** 2
                 for c in c3)
        amp = 1.0 / (q2 + 0.09) ** 2
        for p2 in range(P):
            p4 = ((p2 % ns - c3[0]) % ns) + ns * ((p2 // ns - c3[1]) % ns)
            fv = [0.6, 0.6 if p2 == 0 else 0.3, 0.6 if p3 == 0 else 0.3, 0.6 if p4 == 0 else 0.3]
            expect += w * w * amp * G0 * (fv[2] * fv[3] * (1 - fv[0]) * (1 - fv[1])
                                          - fv[0] * fv[1] * (1 - fv[2]) * (1 - fv[3]))
    assert abs(I1[np.argmax(hot)] - expect) <= 1e-12 * abs(expect)
