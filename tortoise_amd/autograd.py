"""torch.autograd wrapper of the DP5 flow and its hand-written adjoint.

    f_T = DP5Flow.apply(f0, g, omega0, spec)

runs the fixed-step DP5 of the collision operator (tg_dp5_replay) from f0 with
coupling g and phonon energy omega0 (bath n_BE(omega0, T)); backward calls
tg_dp5_adjoint for dL/df0, dL/dg and dL/domega0.  Tensors are CPU float64;
the work runs on the GPU through the C ABI.  spec: dict with keys workload,
steps (accepted step sizes), t0, t1, dense_dt, temperature, ckpt, and
optionally occupation ("nodal" or "hat") and conserve (number-defect removal
in every stage of the forward map, differentiated with it).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import Context


def bath_occupation(omega0, temperature):
    return 1.0 / math.expm1(omega0 / temperature)


def dbath_domega0(omega0, temperature):
    e = math.exp(omega0 / temperature)
    return -e / (temperature * (e - 1.0) ** 2)


def _context(spec, g, omega0):
    w = spec["workload"]
    ctx = Context.create(w, max_batch=1, g=g, omega0=omega0,
                         bath_n=bath_occupation(omega0, spec["temperature"]),
                         occupation=spec.get("occupation", "nodal"))
    ctx.assemble()
    return ctx


class DP5Flow(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f0, g, omega0, spec):
        gv, wv = float(g), float(omega0)
        c = _context(spec, gv, wv)
        f0n = f0.detach().cpu().numpy().astype(np.float64)
        nd = int(round((spec["t1"] - spec["t0"]) / spec["dense_dt"]))
        _, dense = c.dp5_replay(f0n, spec["t0"], spec["t1"], spec["dense_dt"], spec["steps"],
                                max_dense=nd, conserve=spec.get("conserve", False))
        ctx.tg = c
        ctx.spec = spec
        ctx.params = (gv, wv)
        ctx.save_for_backward(f0)
        return torch.from_numpy(dense[-1].copy())

    @staticmethod
    def backward(ctx, grad_out):
        (f0,) = ctx.saved_tensors
        gv, wv = ctx.params
        sp = ctx.spec
        T = sp["temperature"]
        d_f0, d_g, d_w = ctx.tg.dp5_adjoint(
            f0.detach().cpu().numpy(), sp["t0"], sp["t1"], sp["dense_dt"], sp["steps"],
            grad_out.detach().cpu().numpy(), bath_occupation(wv, T), dbath_domega0(wv, T),
            ckpt=sp.get("ckpt", 32), conserve=sp.get("conserve", False))
        return (torch.from_numpy(d_f0), torch.tensor(d_g, dtype=torch.float64),
                torch.tensor(d_w, dtype=torch.float64), None)
