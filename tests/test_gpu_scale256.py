"""A 256x256-section mesh (131072 triangles, 66049 nodes) with the batched
path for 256 members: K is never stored as an n x n array (that alone would
be 35 GB in FP64), only its bands, so the whole context -- both bands, the
projection data, the batched operands and partial sums -- fits in 8 GiB of
device memory.  At this size K's rows still match the oracle's and the
batched evaluation its FP64 one.
"""
import numpy as np
import pytest
import torch  # before the library: the process then has one HIP runtime (torch's)

import oracle
import tortoise_amd as tg
from tortoise_amd import workloads as wl

pytestmark = pytest.mark.gpu

W256 = wl.Workload("square256_ensemble256_mixed", 2, (256, 256), t_final=200.0, rtol=1e-6, atol=1e-12,
                   batch=256)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_square256_context_in_8_gib():
    w = W256
    eps = w.eps()
    centre, amp, temp = wl.ensemble_members(256, seed=0)
    F = np.stack([wl.initial_distribution(eps, w.mu, temp[b], centre[b], w.pump_width, amp[b])
                  for b in range(256)], axis=0)
    bath = np.array([wl.bose_einstein(w.omega0, t) for t in temp])
    # the caller's states and results are not the context's memory
    Fd = torch.from_numpy(F).to("cuda")
    Id = torch.empty_like(Fd)
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    print(f"[gpu] square256: n={w.n_nodes}, assemble FP64 band and the fp16 split, B=256", flush=True)
    ctx = tg.Context.create(w, max_batch=256)
    ctx.assemble(mixed=True)
    ctx.apply_device(Fd.data_ptr(), Id.data_ptr(), batch=256, mixed=True, bath_host=bath)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    used = free0 - free1
    st = ctx.kernel_stats()
    print(f"[gpu] square256: {used / 2**30:.2f} GiB of device memory, nnz(K) = {st['nnz']}", flush=True)
    assert used < 8 * 2**30
    # K rows against the oracle (entries and structural zeros)
    ok = oracle.Kernel(2, w.nsec, w.origin, w.extent, eps, w.g, w.omega0, w.sigma)
    for r in (0, w.n_nodes // 2 + 7, w.n_nodes - 1):
        Kg = ctx.kernel_rows(r, 1)[0]
        Ko = ok.row(r)
        assert np.abs(Kg - Ko).max() <= 1e-12 * np.abs(Ko).max(), r
        assert np.array_equal(Kg != 0, Ko != 0), r
    # the batched evaluation against the FP64 one on a few members
    Im = Id.cpu().numpy()
    pick = [0, 101, 255]
    Ip = ctx.apply(F[pick].T, bath_n=bath[pick])
    for q, b in enumerate(pick):
        assert _rel(Im[b], Ip[:, q]) <= 2e-6, (b, _rel(Im[b], Ip[:, q]))
    ctx.close()
