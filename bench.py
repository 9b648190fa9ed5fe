#!/usr/bin/env python3
"""Scattering-integral throughput on one MI355X.

One step = one pass of the hot path over one batch: B collision-operator
evaluations (B = 1 for the plain FP64 form) on the 128x128-section square
mesh (16641 nodes, dense 16641^2 kernel), inputs resident in HBM.  Prints one
JSON line (see README of the measurement contract in DESIGN.md).
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


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_plain(ctx, torch, f, I, steps, warmup):
    # the library launches on its own stream: record the events there
    stream = torch.cuda.ExternalStream(ctx.stream())
    for _ in range(warmup):
        ctx.apply_device(f.data_ptr(), I.data_ptr())
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        ctx.apply_device(f.data_ptr(), I.data_ptr())
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / steps, wall


def cpu_baseline(ctx, w, budget_s=15.0):
    """Oracle W-form (one pass over K, OpenMP) on the host cores."""
    import numpy as np

    import oracle

    n = ctx.n
    _log("cpu baseline: downloading K")
    K = ctx.kernel_rows(0, n)
    V = ctx.node_volumes()
    f = w.initial_f()
    oracle.collision_apply(K, V, w.bath_n, f)  # warm-up
    times = []
    t_start = time.perf_counter()
    while len(times) < 5 or time.perf_counter() - t_start < budget_s:
        t0 = time.perf_counter()
        oracle.collision_apply(K, V, w.bath_n, f)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": 1.0 / med, "unit": "evals/s", "cores": cores, "kind": "port",
            "sample": f"oracle FP64 W-form apply, n={n}, median of {len(times)} evaluations "
                      f"({sum(times):.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(members shard across GPUs without per-stage communication, see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import tortoise_amd as tg
    from tortoise_amd import workloads as wl

    w = wl.SQUARE128
    torch.cuda.init()
    ctx = tg.Context.create(w, max_batch=1)
    t0 = time.perf_counter()
    ctx.assemble()
    torch.cuda.synchronize()
    _log(f"assembled K (n={ctx.n}) in {time.perf_counter() - t0:.2f} s")
    n = ctx.n
    f = torch.from_numpy(w.initial_f()).to("cuda")
    I = torch.empty_like(f)
    torch.cuda.synchronize()
    ms, wall = time_plain(ctx, torch, f, I, args.steps, args.warmup)
    evals = 1000.0 / ms
    bytes_alg = 8.0 * n * n + 32.0 * n
    gbs = bytes_alg / (ms * 1e-3) / 1e9
    _log(f"plain FP64: {ms:.4f} ms/eval, {gbs:.0f} GB/s algorithmic")
    out = {
        "metric": "scattering-integral evals/s: batched mixed-prec vs FP64 plain @1e-6 rel err",
        "value": evals,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "square128_fp64_plain", "nodes": n, "batch": 1,
                   "mesh": "128x128 sections, 32768 P1 triangles", "form": "plain FP64"},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": None},
        "wall_s_timed": wall,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ctx, w)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
