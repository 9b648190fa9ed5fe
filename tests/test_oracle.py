"""Pins of the CPU oracle against closed forms, invariants and brute force.

Each check is chosen so that a plausible slip in the oracle (a wrong vertex
order, a dropped basis weight, a sign of omega0, a mistyped tableau entry, a
missing controller rule) makes it fail.  No check retypes the oracle's own
formula.
"""
import math

import numpy as np
import pytest

import oracle

PI = math.pi


# ------------------------------------------------------------------ mesh
def test_mesh_2d_reference_vertex_order():
    # Mesh.cpp:55-87: t=0 -> (0,0),(1,0),(1,1); t=1 -> (0,0),(0,1),(1,1)
    conn = oracle.mesh_elements(2, [1, 1])
    assert conn.tolist() == [[0, 1, 3], [0, 2, 3]]
    conn = oracle.mesh_elements(2, [2, 1])
    # element id 2*(sx + sy*nx) + t; node id ix + 3*iy
    assert conn[2].tolist() == [1, 2, 5] and conn[3].tolist() == [1, 4, 5]


def test_mesh_3d_kuhn_order():
    conn = oracle.mesh_elements(3, [1, 1, 1])
    X = oracle.mesh_nodes(3, [1, 1, 1], [0, 0, 0], [1, 1, 1])
    # every simplex walks (0,0,0) -> (1,1,1) switching one axis per step, the
    # axis order being the permutations of (x,y,z) in lexicographic order
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    for t, p in enumerate(perms):
        pts = X[conn[t]]
        assert np.allclose(pts[0], 0) and np.allclose(pts[3], 1)
        for a in range(3):
            step = pts[a + 1] - pts[a]
            assert np.allclose(step, np.eye(3)[p[a]])


@pytest.mark.parametrize("dim,nsec", [(2, [3, 4]), (3, [2, 3, 2])])
def test_simplices_tile_the_box(dim, nsec):
    ext = [1.5, 2.0, 0.5][:dim]
    X = oracle.mesh_nodes(dim, nsec, [0.0] * dim, ext)
    conn = oracle.mesh_elements(dim, nsec)
    vol = oracle.element_volume(dim, nsec, ext)
    vols = []
    for el in conn:
        M = (X[el[1:]] - X[el[0]]).T
        vols.append(abs(np.linalg.det(M)) / math.factorial(dim))
    assert np.allclose(vols, vol, rtol=1e-13)
    assert abs(sum(vols) - np.prod(ext)) < 1e-12
    V = oracle.node_volumes(dim, nsec, ext)
    assert abs(V.sum() - np.prod(ext)) < 1e-12
    # interior node of the 2D mesh touches 6 triangles: V = 2*vol
    if dim == 2:
        assert abs(V[1 + (nsec[0] + 1) * 1] - 6 * vol / 3) < 1e-14


# ------------------------------------------------------ energy projection
def _explicit_triangle_density(e, a, area, x):
    """Level-set form of int phi_a delta(eps - x) over a linear triangle."""
    order = np.argsort(e, kind="stable")
    E = np.asarray(e)[order]
    k = int(np.where(order == a)[0][0])  # a's rank
    e0, e1, e2 = E
    if x <= e0 or x >= e2:
        return 0.0
    if x < e1:
        g = 2 * area * (x - e0) / ((e1 - e0) * (e2 - e0))
        t = (x - e0) / (e1 - e0)
        u = (x - e0) / (e2 - e0)
        lam = [(2 - t - u) / 2, t / 2, u / 2]
    else:
        g = 2 * area * (e2 - x) / ((e2 - e0) * (e2 - e1))
        p = (e2 - x) / (e2 - e1)
        q = (e2 - x) / (e2 - e0)
        lam = [q / 2, p / 2, (2 - p - q) / 2]
    return g * lam[k]


def test_bspline_density_equals_level_set_form():
    rng = np.random.default_rng(0)
    for _ in range(50):
        e = rng.uniform(-1, 1, 3)
        area = rng.uniform(0.1, 2.0)
        for a in range(3):
            for x in rng.uniform(e.min(), e.max(), 7):
                d1 = oracle.simplex_density(2, e, a, area, x)
                d2 = _explicit_triangle_density(e, a, area, x)
                assert abs(d1 - d2) <= 1e-12 * max(1.0, abs(d2))


def test_tet_density_against_sampling():
    """3D: the B-spline density against a histogram of lambda_a over uniform
    samples of the tetrahedron (statistical, 4e6 samples)."""
    rng = np.random.default_rng(1)
    e = np.array([0.1, 0.7, 0.35, 1.2])
    vol = 0.3
    m = 4_000_000
    lam = rng.dirichlet(np.ones(4), m)
    eps = lam @ e
    edges = np.linspace(e.min(), e.max(), 41)
    for a in range(4):
        hist, _ = np.histogram(eps, bins=edges, weights=lam[:, a])
        hist = hist * vol / m
        w2, _ = np.histogram(eps, bins=edges, weights=lam[:, a] ** 2)
        sd = np.sqrt(w2) * vol / m  # standard deviation of each bin estimate
        xs = (np.arange(200) + 0.5) / 200
        integ = []
        for lo, hi in zip(edges[:-1], edges[1:]):
            # 200-point midpoint rule (the density has kinks at the knots)
            vals = [oracle.simplex_density(3, e, a, vol, lo + (hi - lo) * s) for s in xs]
            integ.append((hi - lo) * np.mean(vals))
        integ = np.array(integ)
        assert np.all(np.abs(hist - integ) <= 5 * sd + 1e-12)
        assert abs(hist.sum() - vol / 4) < 1e-3 * vol


def _square(nsec, dim=2):
    origin = [-PI] * dim
    ext = [2 * PI] * dim
    X = oracle.mesh_nodes(dim, nsec, origin, ext)
    return origin, ext, 0.5 * (X ** 2).sum(1), X


@pytest.mark.parametrize("dim,nsec", [(2, [8, 8]), (2, [5, 7]), (3, [3, 3, 3])])
def test_projection_mass_and_first_moment(dim, nsec):
    origin, ext, eps, X = _square(nsec, dim)
    g = oracle.egrid(eps, 0.01 / 8)
    off, ln, R = oracle.projection(dim, nsec, ext, eps, g)
    V = oracle.node_volumes(dim, nsec, ext)
    # sum_b R_ib = int phi_i = V_i  (hats are a partition of unity)
    assert np.abs(R.sum(1) - V).max() <= 1e-12 * V.max()
    # sum_b R_ib E_b = int phi_i eps_P1 = sum_e vol/((d+1)(d+2)) (2 e_i + sum_{a!=i} e_a)
    E = g.e0 + g.de * (off[:, None] + np.arange(R.shape[1])[None, :])
    m1 = (R * E).sum(1)
    conn = oracle.mesh_elements(dim, nsec)
    vol = oracle.element_volume(dim, nsec, ext)
    ref = np.zeros_like(V)
    for el in conn:
        s = eps[el].sum()
        for a in el:
            ref[a] += vol / ((dim + 1) * (dim + 2)) * (s + eps[a])
    assert np.abs(m1 - ref).max() <= 1e-11 * np.abs(ref).max()


def _brute_projection(eps, conn, vol, g, m):
    """R_ib by the centroid rule on an m x m sub-triangulation of every
    triangle (independent of the spline identity)."""
    pts = []
    for i in range(m):
        for j in range(m - i):
            pts.append(((i + 1 / 3) / m, (j + 1 / 3) / m))
            if i + j < m - 1:
                pts.append(((i + 2 / 3) / m, (j + 2 / 3) / m))
    pts = np.array(pts)
    lam = np.stack([1 - pts.sum(1), pts[:, 0], pts[:, 1]], 1)
    wgt = vol / (m * m)
    Rb = np.zeros((eps.size, g.nbins))
    for el in conn:
        ex = lam @ eps[el]
        c = np.floor((ex - g.e0) / g.de).astype(int)
        t = (ex - (g.e0 + c * g.de)) / g.de
        for a in range(3):
            w = wgt * lam[:, a]
            np.add.at(Rb[el[a]], c, w * (1 - t))
            np.add.at(Rb[el[a]], c + 1, w * t)
    return Rb


def test_projection_against_brute_force():
    """The sub-sampled projection converges to the oracle's R (a wrong knot
    set or basis weight would leave an O(1) plateau instead)."""
    nsec = [2, 2]
    origin, ext, eps, X = _square(nsec)
    g = oracle.egrid(eps, 0.2)
    off, ln, R = oracle.projection(2, nsec, ext, eps, g)
    conn = oracle.mesh_elements(2, nsec)
    vol = oracle.element_volume(2, nsec, ext)
    Rd = np.zeros((eps.size, g.nbins))
    for i in range(eps.size):
        Rd[i, off[i]:off[i] + ln[i]] = R[i, : ln[i]]
    errs = [np.abs(Rd - _brute_projection(eps, conn, vol, g, m)).max() / np.abs(Rd).max()
            for m in (100, 200)]
    assert errs[1] < 1e-3 and errs[1] < errs[0] / 1.8


# ------------------------------------------------------------- kernel K
def test_gauss_table_moments():
    de, sigma = 0.01 / 8, 0.01
    mlo, gm = oracle.gauss_table(de, 0.05, sigma, 8.0)
    x = (mlo + np.arange(gm.size)) * de + 0.05
    assert np.all(np.abs(x) <= 8 * sigma + 1e-15)
    # Riemann sums of a Gaussian on a grid of sigma/8 are exact to ~e^-32
    assert abs(gm.sum() * de - 1.0) < 1e-12
    assert abs((gm * x).sum() * de) < 1e-14
    assert abs((gm * x * x).sum() * de - sigma ** 2) < 1e-15


@pytest.fixture(scope="module")
def k8():
    nsec = [8, 8]
    origin, ext, eps, X = _square(nsec)
    kern = oracle.Kernel(2, nsec, origin, ext, eps, g=0.1, omega0=0.05, sigma=0.01)
    return kern, kern.full(), eps


def test_kernel_transpose_is_absorption(k8):
    kern, K, eps = k8
    nsec = kern.nsec
    km = oracle.Kernel(2, nsec, kern.origin, kern.extent, eps, g=0.1, omega0=-0.05, sigma=0.01)
    assert np.abs(km.full() - K.T).max() <= 1e-13 * np.abs(K).max()


def test_kernel_emission_direction_and_amplitude(k8):
    kern, K, eps = k8
    # mean energy transfer of emission: <E_i - E_j>_K = -omega0 up to the
    # P1 smearing (-0.0454, -0.0486, -0.0496 on 8^2, 16^2, 32^2 sections)
    mbar = (kern.R * (kern.grid.e0 + kern.grid.de * (kern.off[:, None]
            + np.arange(kern.R.shape[1])[None, :]))).sum(1) / kern.V
    d = mbar[:, None] - mbar[None, :]
    assert abs((K * d).sum() / K.sum() + 0.05) < 0.006
    # amplitude: K scales with g^2
    k2 = oracle.Kernel(2, kern.nsec, kern.origin, kern.extent, eps, g=0.2, omega0=0.05,
                       sigma=0.01)
    assert np.abs(k2.full() - 4 * K).max() <= 1e-15 * np.abs(K).max() * 4


def test_kernel_row_equals_full(k8):
    kern, K, eps = k8
    for i in (0, 17, 80):
        assert np.array_equal(kern.row(i), K[i])


# ------------------------------------------------------------ collision
def test_two_level_rate_equation():
    """Two nodes, emission only from node 1 to node 0 (K_01 = a):
    dN0/dt = a[(n+1) f1 (1-f0) - n f0 (1-f1)] = -dN1/dt."""
    a, n = 0.3, 0.25
    K = np.array([[0.0, a], [0.0, 0.0]])
    V = np.array([0.5, 2.0])
    f = np.array([0.2, 0.7])
    I = oracle.collision_apply(K, V, n, f)
    flux = a * ((n + 1) * f[1] * (1 - f[0]) - n * f[0] * (1 - f[1]))
    assert abs(I[0] - flux / V[0]) < 1e-16 and abs(I[1] + flux / V[1]) < 1e-16


def test_collision_invariants(k8):
    kern, K, eps = k8
    V = kern.V
    rng = np.random.default_rng(4)
    f = rng.uniform(0, 1, V.size)
    I = oracle.collision_apply(K, V, 0.09, f)
    assert abs(np.dot(V, I)) <= 1e-15 * np.dot(V, np.abs(I))
    assert np.abs(oracle.collision_apply(K, V, 0.09, np.zeros_like(f))).max() == 0.0
    assert np.abs(oracle.collision_apply(K, V, 0.09, np.ones_like(f))).max() < 1e-15
    # elastic symmetric kernel (omega0 = 0, n = 0) leaves any constant fixed
    k0 = oracle.Kernel(2, kern.nsec, kern.origin, kern.extent, eps, 0.1, 0.0, 0.01)
    K0 = k0.full()
    assert np.abs(K0 - K0.T).max() <= 1e-13 * np.abs(K0).max()
    assert np.abs(oracle.collision_apply(K0, V, 0.0, np.full(V.size, 0.3))).max() < 1e-15
    # quadratic in f: I(f+g) + I(f-g) - 2 I(f) = 2 * second-order term, and
    # that term is independent of f (third difference vanishes)
    g = rng.uniform(-0.1, 0.1, V.size)
    d2a = (oracle.collision_apply(K, V, 0.09, f + g) + oracle.collision_apply(K, V, 0.09, f - g)
           - 2 * I)
    f2 = rng.uniform(0, 1, V.size)
    d2b = (oracle.collision_apply(K, V, 0.09, f2 + g) + oracle.collision_apply(K, V, 0.09, f2 - g)
           - 2 * oracle.collision_apply(K, V, 0.09, f2))
    assert np.abs(d2a - d2b).max() <= 1e-13 * np.abs(I).max()
    assert np.abs(d2a).max() > 1e-4 * np.abs(I).max()
    # conservative correction: zero defect, small change
    Ic = oracle.collision_apply(K, V, 0.09, f, conserve=True)
    assert np.abs(Ic - I).max() <= 1e-13 * np.abs(I).max()
    # the W-form conserves number for any K (the in/out terms are one double
    # sum read two ways), also for states partly above 1 where f(1-f) < 0;
    # the correction then leaves the operator unchanged
    fo = f.copy()
    fo[: V.size // 3] = 1.0 + 0.5 * rng.uniform(0, 1, V.size // 3)
    Kp = K * (1.0 + 1e-3 * rng.uniform(-1, 1, K.shape))
    Io = oracle.collision_apply(Kp, V, 0.09, fo)
    Ioc = oracle.collision_apply(Kp, V, 0.09, fo, conserve=True)
    assert abs(np.dot(V, Io)) <= 1e-14 * np.dot(V, np.abs(Io))
    assert np.abs(Ioc - Io).max() <= 1e-13 * np.abs(Io).max()


# ------------------------------------------------------------ error norm
def test_error_norm_closed_forms():
    nsec = [6, 5]
    origin, ext, eps, X = _square(nsec)
    conn = oracle.mesh_elements(2, nsec)
    pop = np.full(eps.size, 0.5)  # min(|f|,|1-f|) = 0.5
    rtol, atol = 1e-3, 1e-6
    scale = rtol * 0.5 + atol
    err = np.full(eps.size, 2e-4)
    assert abs(oracle.relative_error(2, conn, err, pop, -1.0, rtol, atol) - 2e-4 / scale) < 1e-14
    # linear error field r = x/scale: mean of x^2 over [-pi,pi]^2 is pi^2/3
    err = X[:, 0] * 1e-4
    val = oracle.relative_error(2, conn, err, pop, -1.0, rtol, atol)
    assert abs(val - 1e-4 * math.sqrt(PI ** 2 / 3) / scale) < 1e-12 * val
    # bosons use scale = rtol*f + atol (Band.cpp:96-99)
    pop = np.full(eps.size, 2.0)
    err = np.full(eps.size, 1e-5)
    assert abs(oracle.relative_error(2, conn, err, pop, 1.0, rtol, atol)
               - 1e-5 / (rtol * 2 + atol)) < 1e-13


# ------------------------------------------------------------------- DP5
def test_dp5_one_step_stability_polynomial():
    """One accepted step on y' = lam y gives y5 = R(lam h) y0 with the DP5
    stability polynomial R(z) = sum_{k<=5} z^k/k! + z^6/600."""
    lam, h = -1.3, 0.4
    r = oracle.dp5_integrate(lambda t, y: lam * y, lambda e, y: 0.0, [1.0], 0.0, h, h)
    z = lam * h
    R = sum(z ** k / math.factorial(k) for k in range(6)) + z ** 6 / 600.0
    assert r["stats"].attempted == 1
    assert abs(r["y"][0] - R) < 1e-15
    # dense output at s = 1 reproduces y5
    assert abs(r["dense"][0][0] - R) < 1e-15


def test_dp5_embedded_error_is_fifth_order():
    errs = []
    for h in (0.2, 0.1, 0.05):
        rec = []
        oracle.dp5_integrate(lambda t, y: -y, lambda e, y: rec.append(e[0]) or 0.0, [1.0],
                             0.0, h, h)
        errs.append(abs(rec[0]))
    assert 28 < errs[0] / errs[1] < 36 and 29 < errs[1] / errs[2] < 35


def test_dp5_cos_known_answer():
    # DormandPrince54.hpp:83-106: y' = cos t, y(0) = 0 -> sin t
    r = oracle.dp5_integrate(lambda t, y: np.array([math.cos(t)]),
                             lambda e, y: abs(e[0]) / (1e-10 + 1e-10 * abs(y[0])),
                             [0.0], 0.0, 10.0, 0.5)
    ts = r["dense_t"]
    assert len(ts) == 20 and abs(ts[-1] - 10.0) < 1e-12
    assert np.abs(r["dense"][:, 0] - np.sin(ts)).max() < 1e-8


def test_dp5_controller_sequence():
    """Prescribed error norms drive the PI controller through reject, the
    post-reject cap and a normal accept (DormandPrince54.hpp:623-661)."""
    seq = iter([4.0, 0.5, 0.02, 2e-3, 0.9])
    r = oracle.dp5_integrate(lambda t, y: -y, lambda e, y: next(seq), [1.0], 0.0, 3.0, 1.0)
    h = r["h"]
    acc = r["accepted"]
    assert acc.tolist() == [False, True, True, True, True][: len(acc)]
    h1 = 1.0 * min(max(0.98 * 4.0 ** (-0.2), 0.333333), 3.0)            # reject
    s2 = min(max(0.98 * 0.5 ** (-0.14) * 1.0 ** 0.08, 0.333333), 3.0)
    h2 = h1 * min(1.0, s2)                                               # capped after reject
    s3 = min(max(0.98 * 0.02 ** (-0.14) * 0.5 ** 0.08, 0.333333), 3.0)
    h3 = h2 * s3
    assert np.allclose(h[:4], [1.0, h1, h2, h3], rtol=1e-15)


def test_dp5_collision_conserves_number(k8):
    kern, K, eps = k8
    f0 = 1 / (np.exp((eps - 0.5) / 0.02) + 1) + 0.2 * np.exp(-0.5 * ((eps - 0.9) / 0.03) ** 2)
    r = oracle.dp5_collision(2, kern.nsec, K, kern.V, 0.0894, f0, 0.0, 20.0, 0.2, 1e-8, 1e-12,
                             max_dense=100)
    n0 = np.dot(kern.V, f0)
    n1 = np.dot(kern.V, r["dense"][-1])
    assert abs(n1 - n0) <= 1e-13 * n0
    # replay of the accepted steps reproduces the run bit for bit
    rp = oracle.dp5_collision_replay(2, kern.nsec, K, kern.V, 0.0894, f0, 0.0, 20.0, 0.2,
                                     r["h"], r["accepted"])
    assert np.array_equal(rp["dense"], r["dense"]) and np.array_equal(rp["y"], r["y"])


def test_band_form_equals_dense_operator(k8):
    """The CPU W-form over K's band in energy order (rows from the first to
    the last nonzero) gives the dense operator's I for any bath."""
    kern, K, eps = k8
    band = oracle.KernelBand(lambda idx: K[idx], eps, kern.V, block=7)
    assert band.entries >= np.count_nonzero(K)
    rng = np.random.default_rng(1)
    f = rng.uniform(0, 1, eps.size)
    for n in (0.0, 0.3):
        a = oracle.collision_apply(K, kern.V, n, f)
        for nt in (1, 3):
            b = band.apply(n, f, nthreads=nt)
            assert np.abs(a - b).max() <= 1e-13 * np.abs(a).max()
