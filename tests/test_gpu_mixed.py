"""GPU parity of the batched mixed-precision path (fp16 hi+lo split, FP32
MFMA accumulation, FP64 epilogue) against the FP64 paths.

Tolerance: relative L2 error of one evaluation <= 2e-6 per member against
the FP64 oracle (the split keeps ~22 significant bits per operand; FP32
accumulation adds ~1e-7).  Number conservation with TG_CONSERVE: the defect
sum_i V_i I_i is at FP64 roundoff.
"""
import numpy as np
import pytest
import torch  # before the library: the process then has one HIP runtime (torch's)

import oracle
import tortoise_amd as tg
from tortoise_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _members(w, B, seed=0):
    eps = w.eps()
    centre, amp, temp = wl.ensemble_members(B, seed)
    F = np.stack([wl.initial_distribution(eps, w.mu, temp[b], centre[b], w.pump_width, amp[b])
                  for b in range(B)], axis=1)
    bath = np.array([wl.bose_einstein(w.omega0, t) for t in temp])
    return F, bath


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def sq32_mixed():
    w = wl.SQUARE32
    print("[gpu] mixed square32 B=8", flush=True)
    ctx = tg.Context.create(w, max_batch=8)
    ctx.assemble(mixed=True)
    ok = oracle.Kernel(w.dim, w.nsec, w.origin, w.extent, w.eps(), w.g, w.omega0, w.sigma)
    return w, ctx, ok, ok.full()


def test_mixed_matches_fp64(sq32_mixed):
    w, ctx, ok, K = sq32_mixed
    F, bath = _members(w, 8)
    Im = ctx.apply(F, bath_n=bath, mixed=True)
    Ip = ctx.apply(F, bath_n=bath)
    for b in range(8):
        Io = oracle.collision_apply(K, ok.V, bath[b], F[:, b])
        assert _rel(Ip[:, b], Io) <= 1e-13
        assert _rel(Im[:, b], Io) <= 2e-6, (b, _rel(Im[:, b], Io))


def test_mixed_band_nonzero_count(sq32_mixed):
    """The nonzeros of [K; K^T] the library counts per band block for the
    roofline: the band holds every nonzero of K but the tail entries beyond
    the extents of the fp16 split (entries whose split is zero, a few in
    ten thousand here), so the count is 2 nnz(K) (the oracle's count of K's
    nonzeros, whose structural zeros the library's K shares) less those;
    with no operand zero for every member every iteration is computed and the
    launch's count is the total."""
    w, ctx, ok, K = sq32_mixed
    st = ctx.kernel_stats()
    assert st["nnz"] == int(np.count_nonzero(K))
    assert (1.0 - 1e-3) * 2 * st["nnz"] <= st["mixed_nnz"] <= 2 * st["nnz"], (st["mixed_nnz"], 2 * st["nnz"])
    eps = w.eps()
    H = np.stack([1.0 / (np.exp((eps - 3.0) / (1.0 + 0.1 * b)) + 1.0) for b in range(8)], axis=1)
    ctx.apply(H, bath_n=np.full(8, 0.1), mixed=True)
    st = ctx.kernel_stats()
    if st["mixed_skip"]:
        assert st["mixed_last_nnz"] == st["mixed_nnz"] and st["mixed_last_band"] == st["mixed_band"]


def test_mixed_partial_batch_and_conservation(sq32_mixed):
    w, ctx, ok, K = sq32_mixed
    F, bath = _members(w, 3, seed=5)
    Im = ctx.apply(F, bath_n=bath, mixed=True, conserve=True)
    for b in range(3):
        Io = oracle.collision_apply(K, ok.V, bath[b], F[:, b], conserve=True)
        assert _rel(Im[:, b], Io) <= 2e-6
        assert abs(np.dot(ok.V, Im[:, b])) <= 1e-13 * np.dot(ok.V, np.abs(Im[:, b]))


def test_mixed_conservation_with_states_above_one(sq32_mixed):
    """A pump inside the Fermi sea pushes f above 1 where f(1-f) < 0: the
    number-defect removal still zeroes sum V I."""
    w, ctx, ok, K = sq32_mixed
    eps = w.eps()
    F = np.stack([wl.initial_distribution(eps, w.mu, 0.02, 0.25, w.pump_width, 0.5),
                  wl.initial_distribution(eps, w.mu, 0.03, 0.9, w.pump_width, 0.1)], axis=1)
    assert F[:, 0].max() > 1.2
    Im = ctx.apply(F, mixed=True, conserve=True)
    for b in range(2):
        assert abs(np.dot(ok.V, Im[:, b])) <= 1e-13 * np.dot(ok.V, np.abs(Im[:, b]))
        Io = oracle.collision_apply(K, ok.V, w.bath_n, F[:, b], conserve=True)
        assert _rel(Im[:, b], Io) <= 2e-6


def test_mixed_edge_states(sq32_mixed):
    w, ctx, ok, K = sq32_mixed
    F = np.zeros((ctx.n, 2))
    F[:, 1] = 1.0
    Im = ctx.apply(F, mixed=True)
    assert np.abs(Im[:, 0]).max() == 0.0
    # full band: (1-f) = 0 and K^T 1 - K^T f cancels against the split sums up
    # to the FP32 accumulation of K^T f (~1e-7 of the column sum)
    assert np.abs(Im[:, 1] * ok.V).max() <= 1e-6 * np.abs(K).sum(0).max()


def test_mixed_large_mesh_full_batch():
    w = wl.SQUARE128_ENSEMBLE
    print("[gpu] mixed square128 B=256", flush=True)
    ctx = tg.Context.create(w, max_batch=256)
    ctx.assemble(mixed=True)
    F, bath = _members(w, 256, seed=0)
    Im = ctx.apply(F, bath_n=bath, mixed=True)
    pick = [0, 17, 100, 255]
    Ip = ctx.apply(F[:, pick], bath_n=bath[pick])
    # one evaluation at 16641 nodes: FP32 sums of at most 32 products folded
    # into FP64; the operand rounding (~2^-22) is amplified by the in/out
    # cancellation of the operator
    for q, b in enumerate(pick):
        assert _rel(Im[:, b], Ip[:, q]) <= 1e-6, (b, _rel(Im[:, b], Ip[:, q]))


def test_mixed_member_rows_outside_lds(monkeypatch):
    """The energy-order gather and scatter through global memory (member rows
    larger than the LDS) give the same bits as the LDS row path."""
    w = wl.SQUARE32
    F, bath = _members(w, 8, seed=7)
    # one tile configuration for both contexts (the autotuner may pick
    # different ones, which round differently)
    for k, v in [("TG_MIXED_NW", "64"), ("TG_MIXED_WAVES", "4"), ("TG_MIXED_FOLD", "1"),
                 ("TG_MIXED_DEPTH", "2"), ("TG_MIXED_SPLITS", "1")]:
        monkeypatch.setenv(k, v)
    ca = tg.Context.create(w, max_batch=8)
    ca.assemble(mixed=True)
    Ia = ca.apply(F, bath_n=bath, mixed=True, conserve=True)
    Ja = ca.apply(F, bath_n=bath, mixed=True)
    monkeypatch.setenv("TG_MIXED_ROWCACHE", "0")
    cb = tg.Context.create(w, max_batch=8)
    cb.assemble(mixed=True)
    Ib = cb.apply(F, bath_n=bath, mixed=True, conserve=True)
    Jb = cb.apply(F, bath_n=bath, mixed=True)
    assert np.array_equal(Ia, Ib)
    assert np.array_equal(Ja, Jb)


def test_mixed_bath_changes_between_calls(sq32_mixed):
    """The library keeps the last per-member bath on the device and uploads
    only when it changes: a new bath, a smaller batch with the same leading
    values, and a return to the first bath each give the FP64 operator."""
    w, ctx, ok, K = sq32_mixed
    F, bath = _members(w, 8, seed=17)
    for bb, cols in ((bath, 8), (bath * 1.5, 8), (bath * 1.5, 5), (bath, 8), (bath, 3)):
        Im = ctx.apply(F[:, :cols], bath_n=bb[:cols], mixed=True)
        for b in range(cols):
            Io = oracle.collision_apply(K, ok.V, bb[b], F[:, b])
            assert _rel(Im[:, b], Io) <= 2e-6, (cols, b)


# every tile configuration the library carries: (members per workgroup,
# waves, FP64 fold interval, prefetch depth), each with a K-split count
TILE_CONFIGS = [(64, 4, 0, 1, 16), (64, 4, 1, 1, 1), (64, 4, 1, 2, 3), (64, 4, 1, 3, 2),
                (64, 4, 2, 2, 1), (128, 4, 0, 1, 8), (128, 4, 1, 1, 2), (128, 8, 1, 1, 1),
                (128, 8, 1, 2, 2), (128, 8, 1, 3, 1), (128, 8, 2, 2, 3), (256, 4, 0, 1, 8),
                (256, 4, 0, 2, 4), (256, 8, 0, 2, 8)]
# the persistent kernel's configurations: (members per workgroup, fold,
# prefetch depth, waves per workgroup, A staged through LDS)
PERSIST_CONFIGS = [(64, 2, 2, 4, 0), (64, 2, 3, 4, 0), (128, 1, 2, 4, 0), (128, 2, 2, 4, 0),
                   (128, 4, 2, 4, 0), (128, 2, 2, 8, 0), (128, 2, 3, 8, 0),
                   (128, 4, 2, 8, 0), (128, 4, 3, 8, 0), (64, 2, 2, 8, 0), (64, 2, 3, 8, 0),
                   (128, 2, 2, 8, 1), (128, 2, 3, 8, 1), (128, 4, 2, 8, 1), (64, 2, 2, 8, 1),
                   (128, 2, 1, 8, 1)]


def test_mixed_tile_configurations(monkeypatch):
    """Each tile configuration gives the FP64 operator to the same tolerance
    and reports itself through the kernel statistics."""
    w = wl.SQUARE32
    B = 256
    F, bath = _members(w, B, seed=11)
    ref = None
    for nw, waves, fold, depth, ns in TILE_CONFIGS:
        print(f"[gpu] tile config {nw} {waves} {fold} {depth} {ns}", flush=True)
        monkeypatch.setenv("TG_MIXED_NW", str(nw))
        monkeypatch.setenv("TG_MIXED_WAVES", str(waves))
        monkeypatch.setenv("TG_MIXED_FOLD", str(fold))
        monkeypatch.setenv("TG_MIXED_DEPTH", str(depth))
        monkeypatch.setenv("TG_MIXED_SPLITS", str(ns))
        ctx = tg.Context.create(w, max_batch=B)
        ctx.assemble(mixed=True)
        st = ctx.kernel_stats()
        assert (st["mixed_persist"], st["mixed_tile"], st["mixed_waves"], st["mixed_fold"],
                st["mixed_depth"], st["mixed_splits"]) == (0, nw, waves, fold, depth, ns)
        if ref is None:
            ref = ctx.apply(F[:, ::37], bath_n=bath[::37])
        Im = ctx.apply(F, bath_n=bath, mixed=True)
        for q, b in enumerate(range(0, B, 37)):
            assert _rel(Im[:, b], ref[:, q]) <= 2e-6, (nw, waves, fold, depth, ns, b)
        ctx.close()


@pytest.mark.parametrize("mesh", ["square32", "square64"])
def test_mixed_persistent_configurations(monkeypatch, mesh):
    """The persistent kernel (balanced panel schedule) in each configuration
    gives the FP64 operator to the same tolerance, and with whole panels as
    work units (no cut bands), one K-split and the same fold interval it sums
    in the order of the tiled kernel: the two results are equal bit for bit.
    With cut bands (here every band longer than 8 iterations computed as two
    chunks) it gives the FP64 operator to the same tolerance."""
    w = wl.SQUARE32 if mesh == "square32" else wl.Workload("square64", 2, (64, 64))
    B = 256
    F, bath = _members(w, B, seed=13)
    for k in ("TG_MIXED_WAVES", "TG_MIXED_SPLITS", "TG_MIXED_AL", "TG_MIXED_CUT"):
        monkeypatch.delenv(k, raising=False)
    ref = None
    tiled = {}
    for fold in (1, 2):
        print(f"[gpu] {mesh} tiled 128x8 fold {fold}", flush=True)
        monkeypatch.setenv("TG_MIXED_PERSIST", "0")
        monkeypatch.setenv("TG_MIXED_NW", "128")
        monkeypatch.setenv("TG_MIXED_WAVES", "8")
        monkeypatch.setenv("TG_MIXED_FOLD", str(fold))
        monkeypatch.setenv("TG_MIXED_DEPTH", "2")
        monkeypatch.setenv("TG_MIXED_SPLITS", "1")
        ctx = tg.Context.create(w, max_batch=B)
        ctx.assemble(mixed=True)
        tiled[fold] = ctx.apply(F, bath_n=bath, mixed=True)
        if ref is None:
            ref = ctx.apply(F[:, ::37], bath_n=bath[::37])
        ctx.close()
    monkeypatch.delenv("TG_MIXED_SPLITS")
    # cut bands, default tile configuration
    monkeypatch.setenv("TG_MIXED_PERSIST", "1")
    monkeypatch.setenv("TG_MIXED_CUT", "8")
    print(f"[gpu] {mesh} persistent, bands cut at 8 iterations", flush=True)
    ctx = tg.Context.create(w, max_batch=B)
    ctx.assemble(mixed=True)
    assert ctx.kernel_stats()["mixed_cuts"] > 0
    Ic = ctx.apply(F, bath_n=bath, mixed=True)
    for q, b in enumerate(range(0, B, 37)):
        assert _rel(Ic[:, b], ref[:, q]) <= 2e-6, b
    ctx.close()
    monkeypatch.setenv("TG_MIXED_CUT", "0")
    for nw, fold, depth, waves, al in PERSIST_CONFIGS:
        print(f"[gpu] {mesh} persistent {nw} fold {fold} depth {depth} waves {waves} A-in-LDS {al}",
              flush=True)
        monkeypatch.setenv("TG_MIXED_PERSIST", "1")
        monkeypatch.setenv("TG_MIXED_NW", str(nw))
        monkeypatch.setenv("TG_MIXED_FOLD", str(fold))
        monkeypatch.setenv("TG_MIXED_DEPTH", str(depth))
        monkeypatch.setenv("TG_MIXED_WAVES", str(waves))
        monkeypatch.setenv("TG_MIXED_AL", str(al))
        ctx = tg.Context.create(w, max_batch=B)
        ctx.assemble(mixed=True)
        st = ctx.kernel_stats()
        assert (st["mixed_persist"], st["mixed_tile"], st["mixed_fold"], st["mixed_depth"],
                st["mixed_waves"], st["mixed_splits"], st["mixed_cuts"]) == (1 + al, nw, fold, depth, waves, 1, 0)
        Im = ctx.apply(F, bath_n=bath, mixed=True)
        for q, b in enumerate(range(0, B, 37)):
            assert _rel(Im[:, b], ref[:, q]) <= 2e-6, (nw, fold, depth, waves, al, b)
        if fold in tiled:
            assert np.array_equal(Im, tiled[fold]), (nw, fold, depth, waves, al)
        # a partial batch runs the same schedule
        Ip = ctx.apply(F[:, :5], bath_n=bath[:5], mixed=True)
        assert np.array_equal(Ip, Im[:, :5])
        ctx.close()


def test_replay_matches_oracle_replay(sq32_mixed):
    """FP64 replay of a recorded step sequence equals the oracle's replay."""
    w, ctx, ok, K = sq32_mixed
    f0 = w.initial_f()
    fg, dg, st, steps = ctx.dp5(f0, 0.0, 20.0, 0.2, 1e-8, 1e-12, max_dense=100, h_cap=4096)
    fr, dr = ctx.dp5_replay(f0, 0.0, 20.0, 0.2, steps, max_dense=100)
    assert _rel(dr[-1], dg[-1]) <= 1e-15
    ro = oracle.dp5_collision_replay(2, w.nsec, K, ok.V, w.bath_n, f0, 0.0, 20.0, 0.2, steps,
                                     np.ones(len(steps), bool))
    assert _rel(dr[-1], ro["dense"][-1]) <= 1e-12


def test_mixed_dp5_error_vs_fp64_on_same_steps(sq32_mixed):
    """Batched mixed-precision DP5 (per-member adaptive steps, rtol 1e-6) vs
    FP64 replaying each member's accepted steps: relative L2 error of
    f(t_final) <= 1e-6; particle-number drift <= 1e-12 (with TG_CONSERVE)."""
    w, ctx, ok, K = sq32_mixed
    F, bath = _members(w, 8, seed=3)
    print("[gpu] mixed dp5 square32 B=8 t=100", flush=True)
    fm, dm, stats, steps = ctx.dp5(F, 0.0, 100.0, 1.0, 1e-6, 1e-12, bath_n=bath, mixed=True,
                                   conserve=True, max_dense=100, h_cap=4096)
    assert all(s["n_dense"] == 100 and not s["failed"] for s in stats)
    # members take different numbers of steps
    assert len({s["attempted"] for s in stats}) > 1
    fr, dr = ctx.dp5_replay(F, 0.0, 100.0, 1.0, steps, bath_n=bath, max_dense=100)
    # each member of the batched FP64 replay takes its own steps: it equals the
    # replay of that member alone, bit for bit
    for b in (0, 5):
        f1, d1 = ctx.dp5_replay(F[:, b], 0.0, 100.0, 1.0, steps[b], bath_n=bath[b], max_dense=100)
        assert np.array_equal(d1, dr[b]) and np.array_equal(f1, fr[:, b])
    V = ok.V
    for b in range(8):
        assert _rel(dm[b, -1], dr[b, -1]) <= 1e-6, (b, _rel(dm[b, -1], dr[b, -1]))
        n0 = np.dot(V, F[:, b])
        assert abs(np.dot(V, dm[b, -1]) - n0) <= 1e-12 * n0


def test_mixed_dp5_device_states_equal_host_states(sq32_mixed):
    """The same batched DP5 with the states in device memory (TG_PTR_DEVICE,
    integrated in place) gives the host-pointer call's states, statistics and
    steps bit for bit."""
    w, ctx, ok, K = sq32_mixed
    F, bath = _members(w, 8, seed=4)
    print("[gpu] mixed dp5 square32 B=8 t=50, device states", flush=True)
    fm, _, stats, steps = ctx.dp5(F, 0.0, 50.0, 1.0, 1e-6, 1e-12, bath_n=bath, mixed=True, conserve=True,
                                  h_cap=4096)
    X = torch.from_numpy(np.ascontiguousarray(F.T)).to("cuda")
    sd, stepd = ctx.dp5_device(X.data_ptr(), 8, 0.0, 50.0, 1.0, 1e-6, 1e-12, bath_n=bath, mixed=True,
                               conserve=True, h_cap=4096)
    torch.cuda.synchronize()
    assert np.array_equal(X.cpu().numpy().T, fm)
    assert sd == stats
    assert all(np.array_equal(a, b) for a, b in zip(stepd, steps))


def test_mixed_dp5_members_independent_of_the_batch():
    """A member's batched DP5 does not depend on the batch around it: in a
    256-member batch the running members are repacked into the leading
    operand columns every attempt (so a member's column and member tile
    change as others finish, and the GEMM's bins regroup when one tile is
    left), yet members 0, 130 and 255 integrated in a batch of three give the
    same states, statistics and steps bit for bit."""
    w = wl.SQUARE32
    B = 256
    F, bath = _members(w, B, seed=21)
    print("[gpu] mixed dp5 square32 B=256 vs B=3, t=60", flush=True)
    ctx = tg.Context.create(w, max_batch=B)
    ctx.assemble(mixed=True)
    fm, _, stats, steps = ctx.dp5(F, 0.0, 60.0, 0.6, 1e-6, 1e-12, bath_n=bath, mixed=True, conserve=True,
                                  h_cap=4096)
    att = np.array([s["attempted"] for s in stats])
    # members finish at different attempts, so the columns move during the run
    assert att.max() > att.min()
    pick = [0, 130, 255]
    fs, _, st3, steps3 = ctx.dp5(F[:, pick], 0.0, 60.0, 0.6, 1e-6, 1e-12, bath_n=bath[pick], mixed=True,
                                 conserve=True, h_cap=4096)
    for q, b in enumerate(pick):
        assert np.array_equal(fs[:, q], fm[:, b]), b
        assert st3[q] == stats[b], b
        assert np.array_equal(steps3[q], steps[b]), b
    ctx.close()


def test_mixed_zero_operand_panels_left_out_exactly(monkeypatch):
    """Panels whose fp16 operands are zero for every member (the empty top of
    cold distributions, where f is below the split's range) are left out of
    the persistent GEMM and read as 0: the results are bit for bit those with
    every panel computed, on the cold 256-member ensemble (most panels left
    out) and on hot members (none left out), and the batched DP5 takes the
    same steps."""
    w = wl.SQUARE128_ENSEMBLE
    F, bath = _members(w, 256, seed=0)
    eps = w.eps()
    H = np.stack([1.0 / (np.exp((eps - 3.0) / (1.0 + 0.01 * b)) + 1.0) for b in range(64)], axis=1)
    for k, v in [("TG_MIXED_PERSIST", "1"), ("TG_MIXED_NW", "128"), ("TG_MIXED_WAVES", "8"),
                 ("TG_MIXED_FOLD", "2"), ("TG_MIXED_DEPTH", "2"), ("TG_MIXED_AL", "1"),
                 ("TG_MIXED_AUTOTUNE", "0")]:
        monkeypatch.setenv(k, v)
    out = {}
    for skip in ("1", "0"):
        monkeypatch.setenv("TG_MIXED_SKIP", skip)
        print(f"[gpu] mixed square128 B=256, zero-operand panels skipped={skip}", flush=True)
        ctx = tg.Context.create(w, max_batch=256)
        ctx.assemble(mixed=True)
        Ic = ctx.apply(F, bath_n=bath, mixed=True)
        st_cold = ctx.kernel_stats()
        Ih = ctx.apply(H, bath_n=bath[:64], mixed=True)
        st_hot = ctx.kernel_stats()
        _, _, dst = ctx.dp5(F[:, :32], 0.0, 20.0, 0.2, 1e-6, 1e-12, bath_n=bath[:32], mixed=True)
        out[skip] = (Ic, Ih, [s["attempted"] for s in dst], st_cold, st_hot)
        ctx.close()
    a, b = out["1"], out["0"]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]
    assert a[3]["mixed_skip"] == 1 and b[3]["mixed_skip"] == 0
    # the cold ensemble leaves most of the band out, the hot members none of it
    assert a[3]["mixed_last_band"] < 0.6 * a[3]["mixed_band"], a[3]
    assert a[4]["mixed_last_band"] == a[4]["mixed_band"], a[4]
    # the nonzeros of [K; K^T] the launch computed: fewer on the cold
    # ensemble, all of them on the hot members and with every unit computed
    assert a[3]["mixed_last_nnz"] < 0.6 * a[3]["mixed_nnz"]
    assert a[4]["mixed_last_nnz"] == a[4]["mixed_nnz"] and b[3]["mixed_last_nnz"] == b[3]["mixed_nnz"]


def test_mixed_cut_bands(monkeypatch):
    """Bands cut in two work units (here every band longer than 6
    iterations): each chunk is summed on its own into a partial-sum slot and
    the epilogue adds the slots in order, so the result does not depend on
    which chunks the schedule leaves out (bit for bit the same with every
    unit computed), matches the FP64 operator to the mixed tolerance, and the
    batched DP5 takes the same steps with and without leaving units out."""
    w = wl.Workload("square64", 2, (64, 64))
    B = 256
    F, bath = _members(w, B, seed=21)
    for k, v in [("TG_MIXED_PERSIST", "1"), ("TG_MIXED_NW", "128"), ("TG_MIXED_WAVES", "8"),
                 ("TG_MIXED_FOLD", "2"), ("TG_MIXED_DEPTH", "2"), ("TG_MIXED_AL", "1"),
                 ("TG_MIXED_CUT", "6")]:
        monkeypatch.setenv(k, v)
    out = {}
    for skip in ("1", "0"):
        monkeypatch.setenv("TG_MIXED_SKIP", skip)
        print(f"[gpu] mixed square64 B=256, bands cut at 6 iterations, skip={skip}", flush=True)
        ctx = tg.Context.create(w, max_batch=B)
        ctx.assemble(mixed=True)
        st = ctx.kernel_stats()
        I = ctx.apply(F, bath_n=bath, mixed=True, conserve=True)
        st_run = ctx.kernel_stats()
        _, _, dst = ctx.dp5(F[:, :16], 0.0, 10.0, 0.1, 1e-6, 1e-12, bath_n=bath[:16], mixed=True, conserve=True)
        if skip == "1":
            Ip = ctx.apply(F[:, ::51], bath_n=bath[::51], conserve=True)
            for q, b in enumerate(range(0, B, 51)):
                assert _rel(I[:, b], Ip[:, q]) <= 2e-6, b
        out[skip] = (I, [s["attempted"] for s in dst], st, st_run)
        ctx.close()
    a, b = out["1"], out["0"]
    assert a[2]["mixed_cuts"] > 0 and a[2]["mixed_cuts"] == b[2]["mixed_cuts"]
    assert np.array_equal(a[0], b[0]) and a[1] == b[1]
    # the cold ensemble leaves units out
    assert a[3]["mixed_last_band"] < a[3]["mixed_band"], a[3]
