"""tg_collision_apply_noalloc (the evaluation that allocates nothing: its
scratch is one device array per GPU) against tg_collision_apply on the same
device inputs: bitwise equal (the same kernels in the same order), on two
calls in a row, the second with a smaller input (another mesh), and a third
back on the first mesh; refusals of the paths it does not take."""
import numpy as np
import pytest

import tortoise_amd as tg
from tortoise_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _run(ctx, F, bath, conserve):
    import torch

    f = torch.tensor(np.ascontiguousarray(F.T), dtype=torch.float64, device="cuda")
    a = torch.zeros_like(f)
    b = torch.full_like(f, np.nan)
    B = F.shape[1]
    ctx.apply_device(f.data_ptr(), a.data_ptr(), batch=B, conserve=conserve, bath_host=bath)
    torch.cuda.synchronize()
    rc = tg.collision_apply_noalloc(ctx, f.data_ptr(), b.data_ptr(), batch=B, bath_host=bath,
                                    conserve=conserve)
    assert rc == tg.TG_OK, tg.load().tg_last_error(ctx._h).decode()
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


def _members(w, B, seed):
    rng = np.random.default_rng(seed)
    f0 = w.initial_f()
    F = np.stack([np.clip(f0 + 0.05 * rng.standard_normal(f0.size), 0, 1) for _ in range(B)], 1)
    return F, rng.uniform(0.05, 0.2, B)


def test_noalloc_matches_apply_on_two_meshes(gpu_available):
    big = wl.Workload("square32", 2, (32, 32))
    small = wl.Workload("square16", 2, (16, 16))
    print("[gpu] no-alloc evaluation, square32 then square16", flush=True)
    c1 = tg.Context.create(big, max_batch=4)
    c1.assemble()
    c2 = tg.Context.create(small, max_batch=4)
    c2.assemble()
    F1, b1 = _members(big, 3, 0)
    for conserve in (False, True):
        a, b = _run(c1, F1, b1, conserve)
        assert np.array_equal(a, b)
    # second call: a smaller input (fewer nodes, one member)
    F2, b2 = _members(small, 1, 1)
    a, b = _run(c2, F2, b2, True)
    assert np.array_equal(a, b)
    # back on the larger mesh: nothing left in the scratch changes the result
    a, b = _run(c1, F1[:, :2], b1[:2], False)
    assert np.array_equal(a, b)
    c1.close()
    c2.close()


def test_noalloc_hat_occupation_matches_apply(gpu_available):
    """With hat-averaged occupation factors the allocation-free evaluation
    keeps them in its scratch too: bitwise equal to tg_collision_apply, on
    a 3D mesh then a smaller 2D one."""
    cube = wl.Workload("cube6", 3, (6, 6, 6), sigma=0.02)
    small = wl.Workload("square16", 2, (16, 16))
    print("[gpu] no-alloc evaluation with hat factors, cube6 then square16", flush=True)
    c1 = tg.Context.create(cube, max_batch=2, occupation="hat")
    c1.assemble()
    c2 = tg.Context.create(small, max_batch=1, occupation="hat")
    c2.assemble()
    F1, b1 = _members(cube, 2, 3)
    for conserve in (False, True):
        a, b = _run(c1, F1, b1, conserve)
        assert np.array_equal(a, b)
    F2, b2 = _members(small, 1, 4)
    a, b = _run(c2, F2, b2, True)
    assert np.array_equal(a, b)
    c1.close()
    c2.close()


def test_noalloc_refuses_what_it_does_not_do(gpu_available):
    import torch

    w = wl.Workload("square8", 2, (8, 8))
    ctx = tg.Context.create(w, max_batch=1)
    ctx.assemble()
    f = torch.zeros(w.n_nodes, dtype=torch.float64, device="cuda")
    out = torch.zeros_like(f)
    # host pointers, the mixed path and kernel timing are refused
    assert tg.collision_apply_noalloc(ctx, f.data_ptr(), out.data_ptr(), flags=0) == tg.TG_ERR_ARG
    assert tg.collision_apply_noalloc(ctx, f.data_ptr(), out.data_ptr(),
                                      flags=tg.TG_PTR_DEVICE | tg.TG_PREC_MIXED) == tg.TG_ERR_ARG
    assert tg.collision_apply_noalloc(ctx, f.data_ptr(), out.data_ptr(),
                                      flags=tg.TG_PTR_DEVICE | tg.TG_TIME_KERNEL) == tg.TG_ERR_ARG
    ctx.close()
