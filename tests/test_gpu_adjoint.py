"""GPU discrete adjoint of DP5 (gradients wrt g, omega0, f0) against the FP64
oracle adjoint and against central finite differences (FP64, h = 1e-6
relative) on the frozen step sequence.  The large case is the 64x64-section
mesh (4225 nodes) to t = 50 with checkpoints every 32 steps, starting from
g = 0.12, omega0 = 0.045 with the target generated at g = 0.1, omega0 = 0.05."""
import math

import numpy as np
import pytest
import torch

import oracle
import tortoise_amd as tg
from tortoise_amd import autograd as tga
from tortoise_amd import workloads as wl

pytestmark = pytest.mark.gpu
T = 0.02


def _bins(w, V, nb=40):
    eps = w.eps()
    edges = np.linspace(0.0, 2.0, nb + 1)
    Wb = np.zeros((nb, eps.size))
    for b in range(nb):
        sel = (eps >= edges[b]) & (eps < edges[b + 1])
        Wb[b, sel] = V[sel]
    return Wb


def test_adjoint_matches_oracle_adjoint():
    w = wl.Workload("square32_adj", 2, (32, 32), g=0.12, omega0=0.045, t_final=4.0)
    print("[gpu] adjoint vs oracle, square32 t=4", flush=True)
    bath, dbath = tga.bath_occupation(w.omega0, T), tga.dbath_domega0(w.omega0, T)
    ctx = tg.Context.create(w, max_batch=1, bath_n=bath)
    ctx.assemble()
    f0 = w.initial_f()
    _, _, _, steps = ctx.dp5(f0, 0.0, 4.0, 0.04, 1e-8, 1e-12, bath_n=bath, max_dense=100,
                             h_cap=4096)
    rng = np.random.default_rng(0)
    dL = rng.normal(size=ctx.n)
    g_f0, g_g, g_w = ctx.dp5_adjoint(f0, 0.0, 4.0, 0.04, steps, dL, bath, dbath, ckpt=5)
    ok = oracle.Kernel(2, w.nsec, w.origin, w.extent, w.eps(), w.g, w.omega0, w.sigma)
    o_f0, o_g, o_w = oracle.dp5_adjoint(ok.full(), ok.domega_full(), ok.V, bath, dbath, w.g, f0,
                                        0.0, 4.0, 0.04, steps, dL, ckpt=7)
    assert np.linalg.norm(g_f0 - o_f0) <= 1e-10 * np.linalg.norm(o_f0)
    assert abs(g_g - o_g) <= 1e-10 * abs(o_g) and abs(g_w - o_w) <= 1e-10 * abs(o_w)


def test_autograd_gradients_against_central_fd():
    w = wl.SQUARE64
    print("[gpu] adjoint square64 t=50 vs central FD", flush=True)
    # target distribution at the true parameters
    tgt_ctx = tg.Context.create(w, max_batch=1, g=0.1, omega0=0.05,
                                bath_n=tga.bath_occupation(0.05, T))
    tgt_ctx.assemble()
    f0 = w.initial_f()
    _, dt_tgt, _ = tgt_ctx.dp5(f0, 0.0, 50.0, 0.5, 1e-8, 1e-12, max_dense=100)
    V = tgt_ctx.node_volumes()
    Wb = torch.from_numpy(_bins(w, V))
    target = Wb @ torch.from_numpy(dt_tgt[-1])
    # frozen steps of the perturbed start g = 0.12, omega0 = 0.045
    g0, w0 = 0.12, 0.045
    c0 = tg.Context.create(w, max_batch=1, g=g0, omega0=w0, bath_n=tga.bath_occupation(w0, T))
    c0.assemble()
    _, _, st, steps = c0.dp5(f0, 0.0, 50.0, 0.5, 1e-8, 1e-12, max_dense=100, h_cap=4096)
    spec = dict(workload=w, steps=steps, t0=0.0, t1=50.0, dense_dt=0.5, temperature=T, ckpt=32)
    assert st["accepted"] > 32  # more than one checkpoint segment

    def loss_of(f0t, g, om):
        fT = tga.DP5Flow.apply(f0t, g, om, spec)
        d = Wb @ fT - target
        return (d * d).sum()

    f0t = torch.from_numpy(f0.copy()).requires_grad_(True)
    gt = torch.tensor(g0, dtype=torch.float64, requires_grad=True)
    wt = torch.tensor(w0, dtype=torch.float64, requires_grad=True)
    L = loss_of(f0t, gt, wt)
    L.backward()

    def L_at(g, om, f):
        with torch.no_grad():
            return float(loss_of(torch.from_numpy(f), torch.tensor(g, dtype=torch.float64),
                                 torch.tensor(om, dtype=torch.float64)))

    hg, hw = 1e-6 * g0, 1e-6 * w0
    fd_g = (L_at(g0 + hg, w0, f0) - L_at(g0 - hg, w0, f0)) / (2 * hg)
    fd_w = (L_at(g0, w0 + hw, f0) - L_at(g0, w0 - hw, f0)) / (2 * hw)
    assert abs(gt.grad.item() - fd_g) <= 1e-5 * abs(fd_g), (gt.grad.item(), fd_g)
    assert abs(wt.grad.item() - fd_w) <= 1e-5 * abs(fd_w), (wt.grad.item(), fd_w)
    gf = f0t.grad.numpy()
    for i in np.argsort(-np.abs(gf))[:3]:
        e = np.zeros_like(f0)
        e[i] = 1e-6
        fd_i = (L_at(g0, w0, f0 + e) - L_at(g0, w0, f0 - e)) / 2e-6
        assert abs(gf[i] - fd_i) <= 1e-5 * abs(fd_i), (i, gf[i], fd_i)
