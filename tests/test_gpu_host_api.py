"""The C++ host-layer example runs the GPU operator through the Tortoise::
API and agrees with the same integration driven through the C ABI."""
import math
import os
import subprocess

import numpy as np
import pytest

import oracle
import tortoise_amd as tg
from tortoise_amd import workloads as wl

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tortoise_amd", "host", "examples", "thermalisation_square")


def _host_form(w, **kw):
    """A C-ABI context with the host layer's W-form: hat-averaged occupation
    factors, whose DP5 runs with the number-defect removal (TG_CONSERVE)."""
    ctx = tg.Context.create(w, max_batch=1, occupation="hat", **kw)
    ctx.assemble()
    return ctx


def test_host_example_matches_abi():
    assert os.path.exists(EXE), "build with `make -C tortoise_amd/host`"
    print("[gpu] host example 16 sections t=10", flush=True)
    r = subprocess.run([EXE, "16", "10", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().splitlines()
    steps = int(lines[0].split()[1])
    t0_num = float(lines[2].split("|")[1])
    fin = lines[-1].split("|")
    num_end, en_end = float(fin[1]), float(fin[2])
    # particle number conserved by the run
    assert abs(num_end - t0_num) <= 1e-12 * t0_num
    # same integration through the C ABI
    w = wl.Workload("square16", 2, (16, 16), t_final=10.0, rtol=1e-8, atol=1e-12)
    ctx = _host_form(w)
    f, dense, st = ctx.dp5(w.initial_f(), 0.0, 10.0, 0.1, 1e-8, 1e-12, max_dense=100, conserve=True)
    assert st["attempted"] == steps
    V = ctx.node_volumes()
    assert abs(np.dot(V, dense[-1]) - num_end) <= 1e-12 * num_end
    # energy with the exact P1 x P1 integral of Function::integrate(rhs)
    conn = oracle.mesh_elements(2, [16, 16])
    eps = w.eps()
    vol = (2 * np.pi / 16) ** 2 / 2
    a, b = dense[-1][conn], eps[conn]
    en = vol / 12 * ((a * b).sum() + (a.sum(1) * b.sum(1)).sum())
    assert abs(en - en_end) <= 1e-11 * abs(en)


def test_host_example_ee_channel_matches_abi():
    """The 4-leg e-e channel added through addScatteringChannelsWildcards<4>
    reaches the backend: same step count and final number/energy as the C ABI
    with the same e-e parameters, and the e-e channel changes the run."""
    assert os.path.exists(EXE), "build with `make -C tortoise_amd/host`"
    print("[gpu] host example 16 sections t=2 with e-e g=1", flush=True)
    r = subprocess.run([EXE, "16", "2", "1", "1.0"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().splitlines()
    steps = int(lines[0].split()[1])
    fin = lines[-1].split("|")
    num_end = float(fin[1])
    w = wl.Workload("square16ee", 2, (16, 16), t_final=2.0, rtol=1e-8, atol=1e-12,
                    ee_g=1.0, ee_kappa2=0.09, ee_sigma=0.05)
    ctx = _host_form(w)
    f, dense, st = ctx.dp5(w.initial_f(), 0.0, 2.0, 0.02, 1e-8, 1e-12, max_dense=100, conserve=True)
    assert st["attempted"] == steps
    V = ctx.node_volumes()
    assert abs(np.dot(V, dense[-1]) - num_end) <= 1e-12 * num_end
    # without e-e the same run differs
    cp = _host_form(w, ee_g=0.0)
    fp, _, _ = cp.dp5(w.initial_f(), 0.0, 2.0, 0.02, 1e-8, 1e-12, max_dense=100, conserve=True)
    assert np.linalg.norm(fp - f) > 1e-6 * np.linalg.norm(f)


def _second_draw_seed(nmc):
    """The host layer draws MC points from one splitmix64 stream: the rates
    call takes the first draw (3 numbers per point in 2D), the propagate call
    the second, which starts 3 * nmc increments later."""
    return (3 * nmc * 0x9E3779B97F4A7C15) % 2 ** 64


def test_host_example_reference_operator_matches_abi():
    """Material::referenceOperator: MaterialTimeStatus::propagate draws the MC
    points (splitmix64 from seed 0, the reference's simplex folds) and runs
    DP5 with the tuple operator on the DG state; Material::scatteringRates
    gives the initial state's rates.  The same points, tuples and calls
    through the C ABI give the same rates integral, step count and final
    number."""
    assert os.path.exists(EXE), "build with `make -C tortoise_amd/host`"
    print("[gpu] host example 16 sections t=2 reference operator", flush=True)
    r = subprocess.run([EXE, "16", "2", "1", "0", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().splitlines()
    rates_int = float(next(x for x in lines if x.startswith("rates integral")).split()[-1])
    i0 = next(i for i, x in enumerate(lines) if x.startswith("steps"))
    steps = int(lines[i0].split()[1])
    t0_num = float(lines[i0 + 2].split("|")[1])
    num_end = float(lines[-1].split("|")[1])
    en_end = float(lines[-1].split("|")[2])
    assert abs(num_end - t0_num) <= 1e-12 * t0_num
    w = wl.Workload("square16", 2, (16, 16))
    op = tg.TupleOperator.create(w, n_umklapp=1, num_mc=50, max_batch=1)
    op.set_points(oracle.mc_points(2, 3, 50, 0))
    f0 = oracle.mc_nodal_to_dg(2, w.nsec, w.initial_f())
    vol = (2 * np.pi / 16) ** 2 / 2

    def integral(g):
        return vol * (g[:, 0] + (g[:, 1] + g[:, 2]) / 3).sum()
    rr = op.rates(f0, w.bath_n)
    assert abs(integral(rr) - rates_int) <= 1e-12 * abs(rates_int)
    # MaterialTimeStatus::propagate redraws the points (the second draw)
    op.set_points(oracle.mc_points(2, 3, 50, _second_draw_seed(50)))
    f, dense, st = op.dp5(f0, 0.0, 2.0, 0.02, 1e-8, 1e-12, w.bath_n, max_dense=100)
    assert st["attempted"] == steps
    assert abs(integral(dense[-1]) - num_end) <= 1e-12 * num_end
    # the electron energy is not conserved: it checks the dynamics themselves
    conn = oracle.mesh_elements(2, [16, 16])
    eps_dg = oracle.mc_nodal_to_dg(2, w.nsec, w.eps())
    vol_m = np.linalg.inv(oracle.mc_inverse_mass(2, vol))
    en = float(np.einsum("ea,ab,eb->", eps_dg, vol_m, dense[-1]))
    assert abs(en - en_end) <= 1e-11 * abs(en_end)
    op.close()


def test_host_example_dynamical_phonons_matches_abi():
    """Material::dynamicPhonons with the reference operator: the phonon band
    evolves with the electrons.  Same step count, electron number and total
    energy (electrons + phonons) as the C-ABI run; total energy conserved."""
    assert os.path.exists(EXE), "build with `make -C tortoise_amd/host`"
    print("[gpu] host example 12 sections t=2 dynamical phonons", flush=True)
    r = subprocess.run([EXE, "12", "2", "1", "0", "1", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().splitlines()
    i0 = next(i for i, x in enumerate(lines) if x.startswith("steps"))
    steps = int(lines[i0].split()[1])
    e0_el = float(lines[i0 + 2].split("|")[2])
    fin = next(x for x in lines if x.startswith("final")).split("|")
    num_end, en_end = float(fin[1]), float(fin[2])
    ph = next(x for x in lines if x.startswith("phonons")).split()
    ph_num, ph_en = float(ph[1]), float(ph[3])
    w = wl.Workload("square12", 2, (12, 12))
    op = tg.TupleOperator.create(w, n_umklapp=1, num_mc=50, max_batch=1, dynamic=True)
    op.set_points(oracle.mc_points(2, 3, 50, 0))    # the tuple list comes from the first draw
    op.set_points(oracle.mc_points(2, 3, 50, _second_draw_seed(50)))
    fe = oracle.mc_nodal_to_dg(2, w.nsec, w.initial_f())
    fp = np.zeros_like(fe)
    fp[:, 0] = w.bath_n
    f, dense, st = op.dp5(np.stack([fe, fp]), 0.0, 2.0, 0.02, 1e-8, 1e-12, max_dense=100)
    assert st["attempted"] == steps
    vol = (2 * np.pi / 12) ** 2 / 2

    def integral(g):
        return vol * (g[:, 0] + (g[:, 1] + g[:, 2]) / 3).sum()
    last = dense[-1]                 # the dense output at t = 2, the host's last entry
    assert abs(integral(last[0]) - num_end) <= 1e-12 * num_end
    assert abs(integral(last[1]) - ph_num) <= 1e-12 * abs(ph_num)
    # total energy: electrons (exact P1 x P1) plus phonons (omega0 x number)
    assert abs((en_end + ph_en) - (e0_el + 0.05 * integral(fp))) <= 1e-10 * abs(e0_el)
    op.close()


def test_host_example_pulsed_bath_matches_abi():
    """Band::constrained with a time-dependent uniform population: the host
    layer samples constrainedPopulation (64 per dense step) into the device
    bath schedule; the same samples through the C ABI give the same run."""
    assert os.path.exists(EXE), "build with `make -C tortoise_amd/host`"
    print("[gpu] host example 16 sections t=10 pulsed bath", flush=True)
    r = subprocess.run([EXE, "16", "10", "1", "0", "0", "0", "1.5"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().splitlines()
    steps = int(lines[0].split()[1])
    fin = next(x for x in lines if x.startswith("final")).split("|")
    num_end, en_end = float(fin[1]), float(fin[2])
    bath_end = float(next(x for x in lines if x.startswith("bath at the end")).split()[-1])
    w = wl.Workload("square16", 2, (16, 16), t_final=10.0, rtol=1e-8, atol=1e-12)
    # the example's pulse, n_BE + A (1 - x^2)^2 on |x| < 1, sampled with the
    # same IEEE operations (no transcendental beyond expm1 of the constant)
    nbe, tp, width = 1.0 / math.expm1(0.05 / 0.02), 10.0 / 4, 10.0 / 10

    def pulse(t):
        x = (t - tp) / width
        q = 1.0 - x * x
        return nbe + np.where(np.abs(x) < 1.0, 1.5 * q * q, 0.0)

    dt = (10.0 / 100.0) / 64
    ns = int(np.ceil(2.0 * 10.0 / dt)) + 2
    samples = pulse(np.arange(ns) * dt)
    ctx = _host_form(w)
    ctx.set_bath_schedule(samples, 0.0, dt, atol=1e-3, rtol=1e-3)
    f, dense, st = ctx.dp5(w.initial_f(), 0.0, 10.0, 0.1, 1e-8, 1e-12, max_dense=100, schedule=True,
                           conserve=True)
    assert st["attempted"] == steps
    V = ctx.node_volumes()
    assert abs(np.dot(V, dense[-1]) - num_end) <= 1e-12 * num_end
    conn = oracle.mesh_elements(2, [16, 16])
    eps = w.eps()
    vol = (2 * np.pi / 16) ** 2 / 2
    a, b = dense[-1][conn], eps[conn]
    en = vol / 12 * ((a * b).sum() + (a.sum(1) * b.sum(1)).sum())
    assert abs(en - en_end) <= 1e-10 * abs(en)
    # the pulse did change the run, and the last dense state's phonon band is
    # the constraint at t = 10 (enforceValue at the dense time)
    _, dc, _ = ctx.dp5(w.initial_f(), 0.0, 10.0, 0.1, 1e-8, 1e-12, max_dense=100, conserve=True)
    assert np.abs(dc[-1] - dense[-1]).max() > 1e-6
    assert bath_end == pulse(np.array([10.0]))[0]
