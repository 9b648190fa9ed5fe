#!/usr/bin/env python3
"""Scattering-integral throughput on one MI355X.

Workload: square momentum zone, 128x128 sections (32768 P1 triangles, 16641
nodes), e-ph emission kernel of the Gaussian-broadened overlaps, 256
independent members (pump centre, amplitude and temperature drawn with seed
0).  One step = one batched evaluation of the collision operator for all 256
members (fp16 hi+lo split, FP32 MFMA, FP64 epilogue), inputs resident in HBM.
The plain FP64 form (one member per evaluation) is timed in the same run, and
the CPU oracle (FP64, OpenMP) on the host cores.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# the W-form's occupation factors of the headline contexts: the nodal values
# (the lumped form: positivity-preserving, conserving by itself); the
# hat-averaged form is measured beside it (hat_form_line), every evaluation
# with the number-defect removal the DP5 runs use
OCC = "nodal"
FP16_MFMA_PEAK_TFLOPS = 2500.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "mixed_gemm_traffic.json")
TUPLE_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "tuple_propagate_traffic.json")


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_loop(torch, ctx, call, steps, warmup):
    """Average ms per call from HIP events on the library's stream."""
    stream = torch.cuda.ExternalStream(ctx.stream())
    for _ in range(warmup):
        call()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        call()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps, time.perf_counter() - t0


def kernel_ms(ctx, call, reps):
    """Average duration of the dominant kernel (HIP events around it)."""
    tot = 0.0
    for _ in range(reps):
        call()
        tot += ctx.last_kernel_ms()
    return tot / reps


def _cores():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cpu_wform_band(ctx, w, f, bath_n, budget_s=10.0):
    """The W-form on the host cores over K's band in energy order (oracle,
    FP64, OpenMP): the same algorithmic bytes as the GPU's plain evaluation
    (the nonzeros of K once, plus the n-vectors).  One member; median of the
    timed evaluations after one warm-up; parity against the GPU's FP64."""
    import numpy as np

    import oracle

    n = ctx.n
    K = ctx.kernel_rows(0, n)
    band = oracle.KernelBand(lambda idx: K[idx], w.eps(), ctx.node_volumes())
    del K
    hat = oracle.HatStencil(w.dim, w.nsec) if OCC == "hat" else None
    cores = _cores()
    scratch = np.empty((cores + 3) * n)
    ref = band.apply(bath_n, f, nthreads=cores, scratch=scratch, hat=hat, conserve=True)   # warm-up
    times = []
    t_start = time.perf_counter()
    while len(times) < 5 or (time.perf_counter() - t_start < budget_s and len(times) < 200):
        t0 = time.perf_counter()
        band.apply(bath_n, f, nthreads=cores, scratch=scratch, hat=hat, conserve=True)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    gpu = ctx.apply(f, bath_n=bath_n, conserve=True)
    rel = float(np.linalg.norm(gpu - ref) / np.linalg.norm(ref))
    gbs = (8.0 * band.entries + 32.0 * n) / med / 1e9
    _log(f"CPU W-form over K's band ({band.entries} entries): {med * 1e3:.2f} ms/eval on {cores} threads "
         f"({gbs:.1f} GB/s); GPU FP64 vs it {rel:.1e}")
    return {"value": 1.0 / med, "unit": "evals/s", "cores": cores, "kind": "port",
            "band_entries": band.entries, "gbs": gbs, "gpu_fp64_rel_diff": rel,
            "sample": f"oracle FP64 W-form ({OCC} occupation factors, number-defect removal) over K's band "
                      f"(energy order, {band.entries} entries), one member "
                      f"of the 256-member ensemble, n={n}, median of {len(times)} after a warm-up "
                      f"({sum(times):.1f} s)"}


def dp5_batched(ctx, w, F, bath, check=(0, 128, 255)):
    """The batched workload integrated: all members with DP5 to t_final (per-
    member adaptive steps, rtol 1e-6, number-defect removal), the attempt
    graph on the device, states resident in HBM.  Error of f(t_final)
    against FP64 replays of the same accepted steps for a few members;
    particle-number drift for all; energy drift of electrons + the bath's
    FP64 ledger for all."""
    import numpy as np

    import torch

    V = ctx.node_volumes()
    E = V * w.eps()
    n0, e0 = V @ F, E @ F
    # the states are resident in HBM (member-major) before the clock starts;
    # the timed call returns when every member has reached t_final
    X = torch.from_numpy(np.ascontiguousarray(F.T)).to("cuda")
    h_cap = 512
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st, steps = ctx.dp5_device(X.data_ptr(), F.shape[1], 0.0, w.t_final, w.t_final / 100, w.rtol, w.atol,
                               bath_n=bath, mixed=True, conserve=True, h_cap=h_cap)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    fT = X.cpu().numpy().T
    if max(s["accepted"] for s in st) > h_cap:
        raise RuntimeError("step log shorter than a member's accepted steps")
    att = np.array([s["attempted"] for s in st])
    evals = int(sum(s["evals"] for s in st))
    errs = []
    for b in check:
        fr, _ = ctx.dp5_replay(F[:, b], 0.0, w.t_final, w.t_final / 100, steps[b], bath_n=bath[b],
                               conserve=True)
        errs.append(float(np.linalg.norm(fT[:, b] - fr) / np.linalg.norm(fr)))
    drift = float(np.max(np.abs(V @ fT - n0) / np.abs(n0)))
    fmin = float(fT.min())
    ledger = np.array([s["bath_energy"] for s in st])
    edrift = float(np.max(np.abs(E @ fT + ledger - e0) / np.abs(e0)))
    _log(f"dp5 batched: {F.shape[1]} members to t={w.t_final} in {wall:.3f} s, attempts per member "
         f"{att.min()}..{att.max()}, f(t_final) rel err vs FP64 replay {max(errs):.2e}, "
         f"number drift {drift:.1e}, energy drift (electrons + bath ledger) {edrift:.1e}, min f {fmin:.2e}")
    return {"members": int(F.shape[1]), "t_final": w.t_final, "rtol": w.rtol, "wall_s": wall,
            "attempts_min": int(att.min()), "attempts_max": int(att.max()),
            "member_evals_per_s": evals / wall, "f_final_rel_err_vs_fp64": max(errs),
            "checked_members": list(check), "number_drift_rel": drift,
            "energy_drift_rel": edrift, "bath_energy_member0": float(ledger[0]),
            "min_f_t_final": fmin}, fT


def dp5_plain(ctx, w, f0, bath_n):
    """One member in FP64 with DP5 to t_final at rtol 1e-8 (the plain
    configuration on the same mesh)."""
    import numpy as np

    from tortoise_amd import workloads as wl

    p = wl.SQUARE128
    V = ctx.node_volumes()
    E = V * p.eps()
    t0 = time.perf_counter()
    fT, _, st = ctx.dp5(f0, 0.0, p.t_final, p.t_final / 100, p.rtol, p.atol, bath_n=bath_n, conserve=True)
    wall = time.perf_counter() - t0
    drift = float(abs(V @ fT - V @ f0) / abs(V @ f0))
    edrift = float(abs(E @ fT + st["bath_energy"] - E @ f0) / abs(E @ f0))
    _log(f"dp5 plain: t={p.t_final} rtol {p.rtol} in {wall:.3f} s, {st['attempted']} attempts, "
         f"number drift {drift:.1e}, energy drift (electrons + bath ledger) {edrift:.1e}")
    return {"t_final": p.t_final, "rtol": p.rtol, "wall_s": wall, "attempts": st["attempted"],
            "evals_per_s": st["evals"] / wall, "number_drift_rel": drift, "energy_drift_rel": edrift,
            "bath_energy": st["bath_energy"]}


def cube_ee(steps=5, t_final=50.0, adj_t_final=1.0):
    """The 3D workload with both channels: 20^3 sections (48000 tetrahedra,
    9261 nodes, 8000 periodic points), 16 members, e-ph on the batched mixed
    path plus the e-e channel (V(q) = 1/(|q|^2 + 0.09), sigma_ee = 0.02).
    Batched evaluation time, then the ensemble's DP5 run to t_final with
    number drift, and member 0's adjoint gradient wrt g over [0, adj_t_final]."""
    import numpy as np
    import torch

    import tortoise_amd as tg
    from tortoise_amd import workloads as wl

    w = wl.CUBE20
    B = w.batch
    eps = w.eps()
    centre, amp, temp = wl.ensemble_members(B, seed=0)
    F = np.stack([wl.initial_distribution(eps, w.mu, temp[b], centre[b], w.pump_width, amp[b])
                  for b in range(B)], axis=1)
    bath = np.array([wl.bose_einstein(w.omega0, t) for t in temp])
    ctx = tg.Context.create(w, max_batch=B, occupation=OCC)
    ctx.assemble(mixed=True)
    Fd = torch.from_numpy(np.ascontiguousarray(F.T)).to("cuda")
    Id = torch.empty_like(Fd)

    def call():
        ctx.apply_device(Fd.data_ptr(), Id.data_ptr(), batch=B, mixed=True, bath_host=bath, conserve=True)

    ms, _ = time_loop(torch, ctx, call, steps, 1)
    V = ctx.node_volumes()
    t0 = time.perf_counter()
    fT, _, st = ctx.dp5(F, 0.0, t_final, t_final / 100, w.rtol, w.atol, bath_n=bath, mixed=True,
                        conserve=True)
    wall = time.perf_counter() - t0
    drift = float(np.max(np.abs(V @ fT - V @ F) / np.abs(V @ F)))
    att = [s["attempted"] for s in st]
    _log(f"cube20 e-ph + e-e: {B} members in {ms:.2f} ms per batched evaluation; DP5 to "
         f"t={t_final} in {wall:.1f} s ({min(att)}..{max(att)} attempts), number drift {drift:.1e}")
    # gradient of L = |f(t_final) - f0|^2 / 2 wrt g for member 0 (FP64 plain path,
    # e-e included) through the adjoint of its own accepted steps
    from tortoise_amd import autograd as tga
    f0 = F[:, 0].copy()
    t0 = time.perf_counter()
    _, dense, st0, steps = ctx.dp5(f0, 0.0, adj_t_final, adj_t_final / 100, w.rtol, w.atol,
                                   bath_n=bath[0], max_dense=100, h_cap=1 << 16, conserve=True)
    fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, dLdg, _ = ctx.dp5_adjoint(f0, 0.0, adj_t_final, adj_t_final / 100, steps, dense[-1] - f0, bath[0],
                                 tga.dbath_domega0(w.omega0, temp[0]), ckpt=32, conserve=True)
    adj = time.perf_counter() - t0
    ctx.close()
    _log(f"cube20 adjoint, member 0: forward FP64 DP5 to t={adj_t_final} in {fwd:.1f} s "
         f"({st0['accepted']} steps), adjoint in {adj:.1f} s, dL/dg = {dLdg:.6e}")
    return {"workload": w.name, "members": B, "nodes": int(eps.size), "lattice_points": 8000,
            "ms_per_batched_eval": ms, "member_evals_per_s": B * 1000.0 / ms,
            "dp5_t_final": t_final, "dp5_wall_s": wall, "dp5_attempts_max": int(max(att)),
            "number_drift_rel": drift,
            "adjoint_member0": {"loss": "|f(t) - f0|^2 / 2", "t": adj_t_final, "steps": int(st0["accepted"]),
                                "forward_s": fwd, "adjoint_s": adj, "dL_dg": dLdg}}


def hat_form_line(w, F, bath, steps, warmup):
    """The headline workload with the hat-averaged occupation factors
    (TG_OCC_HAT, number-defect removal): batched and plain evaluation rates,
    the 256-member DP5 to t_final (error against FP64 replays, drifts, the
    smallest occupation at t_final: the hat form does not keep f >= 0)."""
    import numpy as np
    import torch

    import tortoise_amd as tg

    B = F.shape[1]
    ctx = tg.Context.create(w, max_batch=B, occupation="hat")
    ctx.assemble(mixed=True)
    Fd = torch.from_numpy(np.ascontiguousarray(F.T)).to("cuda")
    Id = torch.empty_like(Fd)

    def mixed_call():
        ctx.apply_device(Fd.data_ptr(), Id.data_ptr(), batch=B, mixed=True, bath_host=bath, conserve=True)

    def plain_call():
        ctx.apply_device(Fd.data_ptr(), Id.data_ptr(), batch=1, bath_host=bath[:1], conserve=True)

    ms_mixed, _ = time_loop(torch, ctx, mixed_call, steps, warmup)
    ms_plain, _ = time_loop(torch, ctx, plain_call, steps, warmup)
    Im = Id.cpu().numpy().copy()
    pick = [0, 128, 255]
    Ip = ctx.apply(F[:, pick], bath_n=bath[pick], conserve=True)
    rel = max(float(np.linalg.norm(Im[b] - Ip[:, q]) / np.linalg.norm(Ip[:, q])) for q, b in enumerate(pick))
    try:
        d5, _ = dp5_batched(ctx, w, F, bath)
    except tg.TgError as e:
        d5 = {"failed": str(e)}
    ctx.close()
    _log(f"hat form: batched {ms_mixed:.4f} ms ({B * 1000.0 / ms_mixed:.0f} evals/s), plain FP64 "
         f"{ms_plain:.4f} ms, max rel err {rel:.2e}")
    return {"occupation_factors": "hat", "number_defect_removal": True, "members": B,
            "evals_per_s": B * 1000.0 / ms_mixed, "ms_per_step": ms_mixed,
            "plain_fp64_evals_per_s": 1000.0 / ms_plain, "max_rel_err_vs_fp64": rel, "dp5_batched": d5}


def square64_gradients(T=0.02):
    """The differentiation workload: 64x64 sections (4225 nodes), target f(50)
    generated at g = 0.1, omega0 = 0.05; from g = 0.12, omega0 = 0.045 the
    gradients of L = |Wb f(50) - target|^2 (Wb: 40 energy bins of width 0.05,
    node volumes as weights) wrt g, omega0 and f0 through the adjoint of the
    accepted steps (checkpoints every 32), with dL/dg checked against a
    central difference (h = 1e-6 g) of the same fixed-step map; the operator
    is the measured form (occupation factors OCC, number-defect removal in
    every stage, differentiated with it)."""
    import numpy as np

    import tortoise_amd as tg
    from tortoise_amd import autograd as tga
    from tortoise_amd import workloads as wl

    w = wl.SQUARE64
    t1, dt = 50.0, 0.5
    nd = int(round(t1 / dt))

    def ctx_at(g, om):
        c = tg.Context.create(w, max_batch=1, g=g, omega0=om, bath_n=tga.bath_occupation(om, T),
                              occupation=OCC)
        c.assemble()
        return c

    f0 = w.initial_f()
    tgt_ctx = ctx_at(0.1, 0.05)
    _, dense_t, _ = tgt_ctx.dp5(f0, 0.0, t1, dt, 1e-8, 1e-12, max_dense=nd, conserve=True)
    V = tgt_ctx.node_volumes()
    eps = w.eps()
    edges = np.linspace(0.0, 2.0, 41)
    Wb = np.stack([np.where((eps >= a) & (eps < b), V, 0.0) for a, b in zip(edges[:-1], edges[1:])])
    target = Wb @ dense_t[-1]
    tgt_ctx.close()
    g0, w0 = 0.12, 0.045
    ctx = ctx_at(g0, w0)
    t0 = time.perf_counter()
    _, dense, st, steps = ctx.dp5(f0, 0.0, t1, dt, 1e-8, 1e-12, max_dense=nd, h_cap=1 << 14, conserve=True)
    fwd = time.perf_counter() - t0
    d = Wb @ dense[-1] - target
    t0 = time.perf_counter()
    g_f0, g_g, g_w = ctx.dp5_adjoint(f0, 0.0, t1, dt, steps, 2.0 * (Wb.T @ d),
                                     tga.bath_occupation(w0, T), tga.dbath_domega0(w0, T), ckpt=32,
                                     conserve=True)
    adj = time.perf_counter() - t0
    ctx.close()

    def L_at(g):
        c = ctx_at(g, w0)
        _, dn = c.dp5_replay(f0, 0.0, t1, dt, steps, max_dense=nd, conserve=True)
        c.close()
        r = Wb @ dn[-1] - target
        return float(r @ r)

    hg = 1e-6 * g0
    fd_g = (L_at(g0 + hg) - L_at(g0 - hg)) / (2 * hg)
    rel = abs(g_g - fd_g) / abs(fd_g)
    _log(f"square64 gradients: FP64 DP5 to t={t1} in {fwd:.2f} s ({st['accepted']} steps), adjoint "
         f"(g, omega0, f0; checkpoints every 32) in {adj:.2f} s; dL/dg = {g_g:.6e} (central FD "
         f"rel diff {rel:.1e}), dL/domega0 = {g_w:.6e}")
    return {"workload": "square64_gradients", "nodes": int(eps.size), "steps": int(st["accepted"]),
            "forward_s": fwd, "adjoint_s": adj, "dL_dg": g_g, "dL_domega0": g_w,
            "dL_df0_norm": float(np.linalg.norm(g_f0)), "dL_dg_rel_diff_vs_central_fd": rel}


def hat_moments(w, out):
    """<phi_i, I> of a DG result [nel][3] (LF coefficients, inverse mass
    applied) on the nodal hat functions."""
    import numpy as np

    import oracle

    vol = abs(np.prod(np.array(w.extent) / np.array(w.nsec))) / 2
    m = out @ np.linalg.inv(oracle.mc_inverse_mass(2, vol))
    conn = oracle.mesh_elements(2, list(w.nsec))
    H = np.zeros((w.nsec[0] + 1) * (w.nsec[1] + 1))
    np.add.at(H, conn[:, 0], m[:, 0] - m[:, 1] - m[:, 2])
    np.add.at(H, conn[:, 1], m[:, 1])
    np.add.at(H, conn[:, 2], m[:, 2])
    return H


def wform_vs_reference(w, op, f, bath_n, Hws):
    """The W-form's weak form against the reference operator's on one state,
    for each form in Hws (occupation factors -> the W-form's hat moments
    <phi_i, I>): relative L2 of the hat moments and the ratio of the electron
    cooling powers sum_i eps_i <phi_i, I> (W-form / reference)."""
    import numpy as np

    import oracle

    Ht = hat_moments(w, op.propagate(oracle.mc_nodal_to_dg(2, w.nsec, f), bath_n))
    eps = w.eps()
    out = {"cooling_power_reference": float(eps @ Ht)}
    for form, Hw in Hws.items():
        out[form] = {"hat_moments_rel_l2": float(np.linalg.norm(Hw - Ht) / np.linalg.norm(Ht)),
                     "cooling_power_wform": float(eps @ Hw),
                     "cooling_power_ratio": float((eps @ Hw) / (eps @ Ht))}
    return out


def reference_operator(budget_s=12.0, cpu_budget_s=30.0, member0=None):
    """The reference's own operator (tuple list + MC delta-shell integral,
    NMC = 50, umklapp {-1,0,1}^2) on the square zone: 32x32 sections (the
    reference's CPU-runnable case) and 128x128 sections (the mesh of the
    headline workload).  Timed: tuple enumeration, per-tuple moments,
    propagate for 1 and 8 members; the CPU restatement of the reference
    algorithm (per stage: every tuple's MC integral, OpenMP) on the host cores
    at both sizes, with parity of the GPU against it.  member0 = (f, bath,
    W-form hat moments) of the ensemble's member 0 on the 128x128 mesh: the
    W-form's discrepancy from the reference operator on that state."""
    import numpy as np
    import torch

    import oracle
    import tortoise_amd as tg
    from tortoise_amd import workloads as wl

    res = {}
    for ns in (32, 128):
        w = wl.Workload(f"square{ns}", 2, (ns, ns))
        pts = oracle.mc_points(2, 3, 50, 0)
        op = tg.TupleOperator.create(w, max_batch=8)
        t0 = time.perf_counter()
        op.set_points(pts)
        t_first = time.perf_counter() - t0
        t0 = time.perf_counter()
        op.set_points(oracle.mc_points(2, 3, 50, 1))
        t_mom = time.perf_counter() - t0
        T = op.count()
        base = oracle.mc_nodal_to_dg(2, w.nsec, w.initial_f())
        line = {"sections": ns, "elements": op.nel, "tuples": T, "enumerate_and_moments_s": t_first,
                "moments_s": t_mom}
        stream = torch.cuda.ExternalStream(op.stream())
        for B in (1, 8):
            F = torch.tensor(np.stack([base] * B), device="cuda")
            bath = torch.full((B,), w.bath_n, dtype=torch.float64, device="cuda")
            out = torch.empty_like(F)
            torch.cuda.synchronize()

            def call(tk=False):
                op.propagate_device(F.data_ptr(), bath.data_ptr(), out.data_ptr(), B, time_kernel=tk)
            for _ in range(3):
                call()
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            reps = 20 if ns == 32 else 5
            e0.record(stream)
            for _ in range(reps):
                call()
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            kms = 0.0
            for _ in range(reps):
                call(True)
                kms += op.last_kernel_ms()
            kms /= reps
            # compulsory bytes of one launch: both moment lists (27 doubles an
            # entry, two entries a tuple) read once per 8-member chunk, the
            # other-element indices, the states in and out
            byts = T * 2 * (27 * 8 + 4) + 2 * B * op.nel * 3 * 8
            traffic = None
            try:
                with open(TUPLE_TRAFFIC_FILE) as fh:
                    tt = json.load(fh)
                rec = tt["launches"].get(f"square{ns}_b{B}")
                if rec and rec["tuples"] == T:
                    traffic = rec["fetch_kib"] * 1024 * tt["fetch_scale"] + rec["write_kib"] * 1024 * tt["write_scale"]
            except (OSError, KeyError, ValueError):
                traffic = None
            gbs = byts / (kms * 1e-3) / 1e9
            line[f"b{B}"] = {"ms": ms, "evals_per_s": B * 1000.0 / ms, "kernel_ms": kms,
                             "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS,
                                          "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                                          "traffic": traffic, "bytes_per_launch": byts,
                                          "kernel": f"propagate_kernel<2, {1 if B == 1 else 8}>"}}
            if B == 1:
                Fh = F[0].cpu().numpy()
                Ig = out[0].cpu().numpy()
        _log(f"tuple operator {ns}x{ns}: {T} tuples, enumerate+moments {t_first:.2f} s, moments "
             f"{t_mom * 1e3:.1f} ms; propagate B=1 {line['b1']['ms']:.3f} ms "
             f"({line['b1']['roofline']['achieved']:.0f} GB/s), B=8 {line['b8']['ms']:.3f} ms "
             f"({line['b8']['evals_per_s']:.0f} evals/s)")
        # the CPU restatement of the reference algorithm on the GPU-enumerated
        # list: parity and time per evaluation (one warm-up, then the median)
        idx = op.indices()
        fph = np.zeros_like(Fh)
        fph[:, 0] = w.bath_n
        pts_now = oracle.mc_points(2, 3, 50, 1)       # the points the moments hold
        args = (2, w.nsec, w.origin, w.extent, w.eps(), w.omega0, w.g ** 2, 1, pts_now, idx, Fh, fph)
        cores = _cores()
        t0 = time.perf_counter()
        ref = oracle.mc_apply_eph(*args, nthreads=cores)
        warm = time.perf_counter() - t0
        line["max_rel_diff_vs_cpu"] = float(np.abs(Ig - ref).max() / np.abs(ref).max())
        times = []
        limit = budget_s if ns == 32 else cpu_budget_s
        t_start = time.perf_counter()
        while limit > 0 and (len(times) < 3 or (time.perf_counter() - t_start + warm < limit
                                                 and len(times) < 20)):
            t0 = time.perf_counter()
            oracle.mc_apply_eph(*args, nthreads=cores)
            times.append(time.perf_counter() - t0)
        med = float(np.median(times)) if times else warm
        del idx
        line["cpu_reference_algorithm"] = {
            "value": 1.0 / med, "unit": "evals/s", "cores": cores, "kind": "port",
            "sample": f"CPU restatement of ScatteringChannel::propagate (every tuple's MC integral "
                      f"per stage, {T} tuples of the {ns}x{ns}-section mesh, NMC 50), one member, "
                      f"median of {len(times)} after a warm-up ({sum(times) + warm:.1f} s)"}
        line["gpu_vs_cpu_b8"] = line["b8"]["evals_per_s"] * med
        _log(f"  CPU reference algorithm {med:.3f} s/eval on {cores} threads; GPU/CPU "
             f"{line['gpu_vs_cpu_b8']:.0f}x (8 members); max rel diff {line['max_rel_diff_vs_cpu']:.1e}")
        if ns == 128 and member0 is not None:
            f0, b0, Hw = member0
            line["wform_vs_reference_member0"] = wform_vs_reference(w, op, f0, b0, Hw)
            d = line["wform_vs_reference_member0"]
            _log("  W-form vs reference operator, member 0: " + ", ".join(
                f"{k} factors: cooling power ratio {v['cooling_power_ratio']:.3f}, hat moments rel L2 "
                f"{v['hat_moments_rel_l2']:.3f}" for k, v in d.items() if isinstance(v, dict)))
        if ns == 32:
            # the reference's CPU-runnable propagation: DP5 to t = 100 (rtol 1e-8,
            # atol 1e-12, dense step 1), one member, with the tuple operator and
            # with the W-form on the same mesh
            t0 = time.perf_counter()
            _, _, stt = op.dp5(base, 0.0, 100.0, 1.0, 1e-8, 1e-12, bath_n=w.bath_n)
            t_tuple = time.perf_counter() - t0
            cw = tg.Context.create(wl.SQUARE32, max_batch=1, occupation=OCC)
            cw.assemble()
            t0 = time.perf_counter()
            _, _, stw = cw.dp5(wl.SQUARE32.initial_f(), 0.0, 100.0, 1.0, 1e-8, 1e-12, conserve=True)
            t_wform = time.perf_counter() - t0
            # the W-form's discrepancy from the reference operator on the
            # initial state of this configuration, both forms
            f32 = wl.SQUARE32.initial_f()
            Hw32 = {OCC: cw.node_volumes() * cw.apply(f32, conserve=True)}
            cw.close()
            other = "hat" if OCC == "nodal" else "nodal"
            cn = tg.Context.create(wl.SQUARE32, max_batch=1, occupation=other)
            cn.assemble()
            Hw32[other] = cn.node_volumes() * cn.apply(f32, conserve=True)
            cn.close()
            line["wform_vs_reference_initial"] = wform_vs_reference(w, op, f32, w.bath_n, Hw32)
            line["dp5_t100"] = {"tuple_operator": {"wall_s": t_tuple, "attempts": int(stt["attempted"])},
                                "w_form": {"wall_s": t_wform, "attempts": int(stw["attempted"])}}
            _log(f"  DP5 to t=100 (rtol 1e-8): tuple operator {t_tuple:.2f} s ({stt['attempted']} "
                 f"attempts), W-form {t_wform:.3f} s ({stw['attempted']} attempts)")
        op.close()
        res[f"square{ns}"] = line
    return res


def copy_calibration(nbytes):
    """One device-to-device copy of nbytes (the known-size reference for the
    FETCH_SIZE / WRITE_SIZE counters)."""
    import torch

    x = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes}), flush=True)


def gemm_kernel_name(st):
    """Template name of the batched GEMM kernel the tile configuration runs."""
    if st["mixed_persist"]:
        return (f"mixed_gemm_persist_kernel<{st['mixed_tile']}, {st['mixed_fold']}, {st['mixed_depth']}, "
                f"{st['mixed_waves'] // 4}, {'true' if st['mixed_persist'] == 2 else 'false'}>")
    return (f"mixed_gemm_kernel<{st['mixed_tile']}, {st['mixed_waves']}, {st['mixed_fold']}, "
            f"{st['mixed_depth']}>")


def measured_traffic(st):
    """HBM bytes per launch of the batched GEMM from the committed counter
    runs, when they were taken on the tile configuration in use."""
    if not os.path.exists(TRAFFIC_FILE):
        return None, "no counter profile"
    with open(TRAFFIC_FILE) as fh:
        t = json.load(fh)
    cfg = [st["mixed_persist"], st["mixed_tile"], st["mixed_waves"], st["mixed_fold"], st["mixed_depth"],
           st["mixed_splits"], st["mixed_skip"]]
    if t.get("config") != cfg:
        return None, f"counter profile taken on tile configuration {t.get('config')}, running {cfg}"
    return t["fetch_bytes"] * t["fetch_scale"] + t["write_bytes"] * t["write_scale"], "measured"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-dp5", action="store_true", help="skip the full DP5 runs")
    ap.add_argument("--no-all-panels", action="store_true",
                    help="skip the all-panels context (profiling runs of the headline)")
    ap.add_argument("--no-reference-op", action="store_true",
                    help="skip the reference tuple operator measurements")
    ap.add_argument("--only-reference-op", action="store_true",
                    help="only the reference tuple operator measurements (profiling runs)")
    ap.add_argument("--only-ee", action="store_true",
                    help="only the 20^3 e-ph + e-e ensemble line (profiling runs)")
    ap.add_argument("--only-dp5", action="store_true",
                    help="only the assembly and the DP5 runs on the 128x128 ensemble (profiling runs)")
    ap.add_argument("--only-gradients", action="store_true",
                    help="only the 64x64 gradient line (adjoint of DP5)")
    ap.add_argument("--ee-t-final", type=float, default=50.0,
                    help="DP5 end time of the 20^3 e-e ensemble run")
    ap.add_argument("--ee-adjoint-t-final", type=float, default=1.0,
                    help="end time of the 20^3 member-0 adjoint gradient")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device copy of BYTES (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(ensemble members would shard across GPUs with no per-stage communication)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import tortoise_amd as tg
    from tortoise_amd import workloads as wl

    if args.only_reference_op:
        torch.cuda.init()
        print(json.dumps({"reference_operator": reference_operator()}), flush=True)
        return
    if args.only_gradients:
        torch.cuda.init()
        print(json.dumps({"square64_gradients": square64_gradients()}), flush=True)
        return
    if args.only_ee:
        torch.cuda.init()
        print(json.dumps({"cube20_ee": cube_ee(t_final=args.ee_t_final, adj_t_final=args.ee_adjoint_t_final)}),
              flush=True)
        return

    w = wl.SQUARE128_ENSEMBLE
    B = w.batch
    torch.cuda.init()
    ctx = tg.Context.create(w, max_batch=B, occupation=OCC)
    t0 = time.perf_counter()
    ctx.assemble(mixed=True)
    torch.cuda.synchronize()
    st = ctx.kernel_stats()
    n = ctx.n
    _log(f"assembled K and the split [K;K^T] (n={n}) in {time.perf_counter() - t0:.2f} s; "
         f"nnz(K)={st['nnz']} ({st['nnz'] / n / n:.3f} dense), streamed band "
         f"{st['mixed_band']} entries, splits {st['mixed_splits']}")

    eps = w.eps()
    centre, amp, temp = wl.ensemble_members(B, seed=0)
    F = np.stack([wl.initial_distribution(eps, w.mu, temp[b], centre[b], w.pump_width, amp[b])
                  for b in range(B)], axis=0)  # member-major
    bath = np.array([wl.bose_einstein(w.omega0, t) for t in temp])
    Fd = torch.from_numpy(F).to("cuda")
    Id = torch.empty_like(Fd)
    torch.cuda.synchronize()
    if args.only_dp5:
        Fh = F.T.copy()
        res = {"dp5_batched": dp5_batched(ctx, w, Fh, bath)[0],
               "dp5_plain": dp5_plain(ctx, w, Fh[:, 0].copy(), bath[0])}
        ctx.close()
        print(json.dumps(res), flush=True)
        return

    def mixed_call(time_kernel=False):
        ctx.apply_device(Fd.data_ptr(), Id.data_ptr(), batch=B, mixed=True, bath_host=bath,
                         time_kernel=time_kernel, conserve=True)

    def plain_call(time_kernel=False):
        ctx.apply_device(Fd.data_ptr(), Id.data_ptr(), batch=1, bath_host=bath[:1],
                         time_kernel=time_kernel, conserve=True)

    ms_mixed, wall = time_loop(torch, ctx, mixed_call, args.steps, args.warmup)
    ms_gemm = kernel_ms(ctx, lambda: mixed_call(True), 10)
    ms_plain, _ = time_loop(torch, ctx, plain_call, args.steps, args.warmup)
    ms_dual = kernel_ms(ctx, lambda: plain_call(True), 10)

    # accuracy of the measured batched evaluation against the FP64 form
    Im = Id.cpu().numpy().copy()
    Fh = F.T.copy()
    pick = [0, 63, 128, 255]
    Ip = ctx.apply(Fh[:, pick], bath_n=bath[pick], conserve=True)
    rel = max(float(np.linalg.norm(Im[b] - Ip[:, q]) / np.linalg.norm(Ip[:, q]))
              for q, b in enumerate(pick))

    evals_mixed = B * 1000.0 / ms_mixed
    evals_plain = 1000.0 / ms_plain
    # dominant kernel: the MFMA GEMM.  Algorithmic work of one launch = the
    # two products K F and K^T F over the nonzeros of K for B members, two
    # flops per nonzero of [K; K^T] and member: 4 nnz B with every panel
    # computed.  The launch leaves out the iterations whose fp16 operands are
    # zero for every member (their products are exactly 0); the work counted
    # is exact: two flops per member for each nonzero of [K; K^T] in the
    # iterations it computes (mixed_last_nnz).  Dense-equivalent
    # (4 n^2 B, zeros counted) and issued (three fp16 products over the
    # computed band) rates beside it.
    st_run = ctx.kernel_stats()
    exec_frac = st_run["mixed_last_band"] / st["mixed_band"] if st["mixed_band"] else 1.0
    flops_alg = 2.0 * st_run["mixed_last_nnz"] * B
    tflops = flops_alg / (ms_gemm * 1e-3) / 1e12
    dense_tflops = 4.0 * n * n * B / (ms_gemm * 1e-3) / 1e12
    issued = 3 * 2.0 * st_run["mixed_last_band"] * st["mixed_width"] / (ms_gemm * 1e-3) / 1e12
    traffic, traffic_note = measured_traffic(st)
    # plain FP64 pass: algorithmic bytes = the nonzeros of K once (8 nnz) plus
    # f in, I out and two n-vectors (32 n); streamed = the stored band
    gbs_plain = (8.0 * st["nnz"] + 32.0 * n) / (ms_dual * 1e-3) / 1e9
    gbs_streamed = (8.0 * st["plain_band"] + 32.0 * n) / (ms_dual * 1e-3) / 1e9
    _log(f"mixed: {ms_mixed:.4f} ms/batch ({evals_mixed:.0f} evals/s), GEMM {ms_gemm:.4f} ms "
         f"= {tflops:.1f} TFLOP/s algorithmic, {issued:.1f} TFLOP/s issued ({exec_frac:.0%} of the band "
         f"computed); max rel err {rel:.2e}; "
         f"kernel {gemm_kernel_name(st)} splits {st['mixed_splits']}")
    _log(f"plain FP64: {ms_plain:.4f} ms/eval ({evals_plain:.0f} evals/s), K-band pass "
         f"{ms_dual:.4f} ms = {gbs_plain:.0f} GB/s algorithmic, {gbs_streamed:.0f} GB/s streamed")

    out = {
        "metric": "scattering-integral evals/s: batched mixed-prec vs FP64 plain @1e-6 rel err",
        "value": evals_mixed,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_mixed,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16x2-split/f32-acc/f64-epilogue",
        "data": "synthetic",
        "config": {"workload": w.name, "nodes": n, "members": B,
                   "mesh": "128x128 sections, 32768 P1 triangles",
                   "occupation_factors": OCC, "number_defect_removal": True,
                   "kernel_nnz": st["nnz"], "streamed_band": st["mixed_band"],
                   "splits": st["mixed_splits"]},
        "roofline": {"bound": "mfma", "achieved": tflops, "peak": FP16_MFMA_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": tflops / FP16_MFMA_PEAK_TFLOPS,
                     "traffic": traffic, "traffic_note": traffic_note,
                     "kernel": gemm_kernel_name(st), "kernel_ms": ms_gemm,
                     "flops_per_launch": flops_alg, "executed_band_fraction": exec_frac,
                     "executed_nonzeros": st_run["mixed_last_nnz"], "band_nonzeros": st["mixed_nnz"],
                     "flops_per_launch_all_panels": 2.0 * st["mixed_nnz"] * B, "issued_tflops": issued,
                     "issued_frac": issued / FP16_MFMA_PEAK_TFLOPS,
                     "dense_equiv_tflops": dense_tflops,
                     "tile": {"persistent": st["mixed_persist"] > 0, "a_in_lds": st["mixed_persist"] == 2,
                              "members": st["mixed_tile"], "waves": st["mixed_waves"],
                              "fold": st["mixed_fold"], "depth": st["mixed_depth"],
                              "splits": st["mixed_splits"], "autotuned_ms": st["mixed_tuned_ms"]}},
        "plain_fp64": {"evals_per_s": evals_plain, "ms_per_eval": ms_plain,
                       "kernel": "stack_gemv_kernel", "kernel_ms": ms_dual,
                       "achieved_gbs": gbs_plain, "frac_hbm": gbs_plain / HBM_PEAK_GBS,
                       "streamed_gbs": gbs_streamed, "bytes_per_eval": 8.0 * st["nnz"] + 32.0 * n},
        "mixed_vs_plain": evals_mixed / evals_plain,
        "max_rel_err_vs_fp64": rel,
        "wall_s_timed": wall,
    }
    out["value_note"] = ("the t = 0 ensemble; panels and iterations whose fp16 operands are zero for every "
                         "member are left out of the GEMM (exactly: their products are 0); value_all_panels "
                         "and value_t_final are the same evaluation with every panel computed and on the "
                         "ensemble's state at t_final; roofline.achieved counts two flops per member for "
                         "each nonzero of [K; K^T] in the iterations computed")
    # the same batched evaluation with every panel computed (a context built
    # with TG_MIXED_SKIP=0): the state-independent rate
    if not args.no_all_panels:
        os.environ["TG_MIXED_SKIP"] = "0"
        try:
            ctx_all = tg.Context.create(w, max_batch=B, occupation=OCC)
            ctx_all.assemble(mixed=True)
        finally:
            del os.environ["TG_MIXED_SKIP"]

        def all_call(time_kernel=False):
            ctx_all.apply_device(Fd.data_ptr(), Id.data_ptr(), batch=B, mixed=True, bath_host=bath,
                                 time_kernel=time_kernel, conserve=True)

        ms_all, _ = time_loop(torch, ctx_all, all_call, args.steps, args.warmup)
        ms_gemm_all = kernel_ms(ctx_all, lambda: all_call(True), 10)
        ctx_all.close()
        out["value_all_panels"] = B * 1000.0 / ms_all
        out["all_panels"] = {"ms_per_step": ms_all, "gemm_ms": ms_gemm_all,
                             "gemm_tflops": 2.0 * st["mixed_nnz"] * B / (ms_gemm_all * 1e-3) / 1e12}
        _log(f"all panels: {ms_all:.4f} ms/batch ({out['value_all_panels']:.0f} evals/s), "
             f"GEMM {ms_gemm_all:.4f} ms")
    # W-form hat moments of member 0 (FP64), with the measured (hat) and the
    # nodal occupation factors, for the comparison with the reference
    # operator at this mesh
    other = "hat" if OCC == "nodal" else "nodal"
    I0 = ctx.apply(Fh[:, 0], bath_n=bath[0], conserve=True)
    cn = tg.Context.create(w, max_batch=1, occupation=other)
    cn.assemble()
    I0n = cn.apply(Fh[:, 0], bath_n=bath[0], conserve=True)
    cn.close()
    V0 = ctx.node_volumes()
    member0 = (Fh[:, 0].copy(), float(bath[0]), {OCC: V0 * I0, other: V0 * I0n})
    if not args.no_dp5:
        out["dp5_batched"], fT = dp5_batched(ctx, w, Fh, bath)
        out["dp5_plain"] = dp5_plain(ctx, w, Fh[:, 0].copy(), bath[0])
        # the headline evaluation on the ensemble's state at t_final
        Fd.copy_(torch.from_numpy(np.ascontiguousarray(fT.T)))
        torch.cuda.synchronize()
        ms_tf, _ = time_loop(torch, ctx, mixed_call, args.steps, args.warmup)
        st_tf = ctx.kernel_stats()
        out["value_t_final"] = B * 1000.0 / ms_tf
        out["t_final_state"] = {"ms_per_step": ms_tf, "t": w.t_final,
                                "executed_band_fraction": st_tf["mixed_last_band"] / st["mixed_band"]}
        _log(f"state at t={w.t_final}: {ms_tf:.4f} ms/batch ({out['value_t_final']:.0f} evals/s), "
             f"{out['t_final_state']['executed_band_fraction']:.0%} of the band computed")
    if not args.no_dp5:
        out["hat_form"] = hat_form_line(w, Fh, bath, args.steps, args.warmup)
    if not args.no_cpu:
        out["cpu_wform_band"] = cpu_wform_band(ctx, w, Fh[:, 0].copy(), float(bath[0]))
        out["mixed_vs_cpu_wform_band"] = evals_mixed / out["cpu_wform_band"]["value"]
        out["cpu_baseline"] = out["cpu_wform_band"]
    ctx.close()
    if not args.no_dp5:
        out["cube20_ee"] = cube_ee(t_final=args.ee_t_final, adj_t_final=args.ee_adjoint_t_final)
        out["square64_gradients"] = square64_gradients()
    if not args.no_reference_op:
        ro = reference_operator(cpu_budget_s=0.0 if args.no_cpu else 30.0, member0=member0)
        out["reference_operator"] = ro
        if not args.no_cpu:
            # the reference CPU path on the headline mesh: the reference's own
            # algorithm (tuple list, MC integral per tuple) restated in C
            out["cpu_baseline"] = ro["square128"]["cpu_reference_algorithm"]
            out["mixed_vs_cpu"] = evals_mixed / out["cpu_baseline"]["value"]
    if "mixed_vs_cpu" not in out and "cpu_baseline" in out:
        out["mixed_vs_cpu"] = evals_mixed / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
