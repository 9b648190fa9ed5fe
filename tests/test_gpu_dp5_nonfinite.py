"""DP5 step control when a trial step's error norm is not finite
(tortoise_amd/csrc/tg_dp5_kernels.hpp, controller_kernel).

A non-finite norm rejects the step with the smallest scale 1/3, as
DormandPrince54.hpp:646-649 does for an infinite norm; a member fails only
when that drives its step below dense_dt/1e4, and the batch then reports
TG_ERR_NONFINITE.  A member whose state holds a NaN can never produce a
finite norm, so it must fail after exactly ceil(log_3 1e4) = 9 rejections
while the other members of the batch run to the end unaffected.
"""
import math

import numpy as np
import pytest

import tortoise_amd as tg
from tortoise_amd import workloads as wl

pytestmark = pytest.mark.gpu


def test_nan_member_fails_alone():
    w = wl.SQUARE32
    B = 2
    eps = w.eps()
    centre, amp, temp = wl.ensemble_members(B, 3)
    F = np.stack([wl.initial_distribution(eps, w.mu, temp[b], centre[b], w.pump_width, amp[b])
                  for b in range(B)], axis=1)
    bath = np.array([wl.bose_einstein(w.omega0, t) for t in temp])
    F[17, 0] = np.nan
    ctx = tg.Context.create(w, max_batch=B)
    ctx.assemble(mixed=True)
    t1, dt = 20.0, 0.2
    print("[gpu] dp5 B=2 with a NaN member, mixed", flush=True)
    fb, _, st = ctx.dp5(F, 0.0, t1, dt, 1e-6, 1e-12, bath_n=bath, mixed=True, raise_on_fail=False)
    rejections = math.ceil(math.log(1e4) / math.log(3.0))
    assert st[0]["failed"] and st[0]["accepted"] == 0
    assert st[0]["rejected"] == rejections and st[0]["attempted"] == rejections
    assert not st[1]["failed"] and st[1]["accepted"] > 0
    assert np.all(np.isfinite(fb[:, 1]))
    with pytest.raises(tg.TgError, match="non-finite"):
        ctx.dp5(F, 0.0, t1, dt, 1e-6, 1e-12, bath_n=bath, mixed=True)
    print("[gpu] dp5 B=1, the finite member alone", flush=True)
    f1, _, s1 = ctx.dp5(F[:, 1:2], 0.0, t1, dt, 1e-6, 1e-12, bath_n=bath[1:2], mixed=True)
    assert s1[0]["accepted"] == st[1]["accepted"] and s1[0]["rejected"] == st[1]["rejected"]
    assert np.allclose(f1[:, 0], fb[:, 1], rtol=1e-13, atol=1e-15)


def test_overflowing_first_step_recovers():
    """On the 20^3 cube with the e-e channel (quartic in f) a first trial
    step of 5 overflows inside its stages (the stage values reach inf by the
    third evaluation); the controller rejects it and the ones after with
    the smallest scale until the steps are stable, and the run completes
    with the particle number conserved."""
    w = wl.CUBE20
    eps = w.eps()
    centre, amp, temp = wl.ensemble_members(1, 0)
    f0 = wl.initial_distribution(eps, w.mu, temp[0], centre[0], w.pump_width, amp[0])
    bath = np.array([wl.bose_einstein(w.omega0, temp[0])])
    ctx = tg.Context.create(w, max_batch=1)
    ctx.assemble(mixed=True)
    V = ctx.node_volumes()
    print("[gpu] dp5 cube20 e-e with a first step of 5", flush=True)
    f, _, st = ctx.dp5(f0, 0.0, 5.0, 5.0, 1e-6, 1e-12, bath_n=bath, mixed=True, conserve=True)
    assert not st["failed"] and st["rejected"] >= 3 and st["accepted"] > 0
    assert np.all(np.isfinite(f))
    assert abs(V @ f - V @ f0) <= 1e-12 * (V @ f0)
