// tg_dp5.hip -- adaptive Dormand-Prince 5(4) over a batch of members.
//
// Restates Algorithms::dormandPrince54 (vetoer form,
// include/Generics/Algorithms/DormandPrince54.hpp:479-665) for B independent
// members that share the collision operator: every member keeps its own time,
// step size, PI-controller memory and reject flag; an attempted step runs the
// six stage evaluations for all unfinished members at once (one batched
// operator call per stage), then each member is accepted or rejected on its
// own error norm (Band::relativeError, Band.cpp:87-101).  The phonon band is
// a bath: constant, or constrained to a per-member occupation n_b(t)
// (tg_set_bath_schedule), which enforceValue writes at every stage time
// (MaterialStatus::applyConstraints, MaterialStatus.cpp:198-206) and the
// step vetoer reads (MaterialTimeStatus::timeStepVetoer,
// MaterialTimeStatus.cpp:60-100).  The e-ph epilogue of every stage returns
// the member's electron energy flux; the controller integrates it with the
// step weights into the bath's energy ledger.
//
// The whole attempt runs on the device and is captured once as a HIP graph:
//   prep_kernel        members still running take their step size (0 for the
//                      others, whose stages then reproduce y: masked)
//   6 x (stage_kernel, operator evaluation), y5 and k7
//   error_kernel       err = y5 - y4;  relerr sums (Band::relativeError)
//   controller_kernel  per member: accept/reject, PI update, post-reject cap,
//                      step underflow, dense-output range, step log
//   dense_kernel       dense outputs of the accepted steps
//   accept_kernel      y <- y5, k1 <- k7 (FSAL)
//   a copy of the "any member running" flag to pinned host memory.
// On the batched mixed-precision path (member rows in LDS, no e-e channel)
// the stage, pack, epilogue and error work between two GEMMs is one fused
// member pass (Dp5PassArgs): stage 2 + pack, GEMM, then five times finish
// k_s + next stage + pack, GEMM, and finish k7 + err + the error-norm sums;
// the attempt is then 13 kernels and one schedule per GEMM instead of 37.
// The host replays the graph one attempt ahead of the flag it reads, so it
// never waits for a decision and at most one masked attempt runs at the end.
#include <cmath>
#include <cstring>
#include <vector>

#include "tg_dp5_kernels.hpp"
#include "tg_internal.hpp"

namespace tg {

struct Dp5Work {
    int32_t cap = 0;          // members allocated
    int64_t n = 0;
    double* k[7] = {};        // k1..k7
    double* ys = nullptr;
    double* y5 = nullptr;
    double* err = nullptr;
    double* y = nullptr;
    double* h = nullptr;      // device per-member step
    int8_t* acc = nullptr;
    MemberDev* md = nullptr;
    double* norms = nullptr;  // error-norm sums per member
    double* npart = nullptr;  // their block partials
    double* sched = nullptr;  // replay schedule (B x h_cap)
    int32_t* nsched = nullptr;
    size_t sched_cap = 0;
    double* hlog = nullptr;   // accepted step sizes (B x h_cap)
    size_t hlog_cap = 0;
    double* flux = nullptr;   // e-ph electron energy flux per stage (k1..k7) and member: 7 x cap
    int32_t* running = nullptr;        // device flag
    int32_t* running_host = nullptr;   // pinned copy
    double* dense = nullptr;  // device dense buffer when the caller's is host memory
    size_t dense_cap = 0;
};

static Dp5Work* work_of(tg_ctx* c) { return (Dp5Work*)c->dp5; }

static tg_status dp5_ensure(tg_ctx* c, int32_t B) {
    Dp5Work* w = work_of(c);
    if (w && w->cap >= B) return TG_OK;
    if (w) dp5_destroy(c);
    w = new Dp5Work();
    c->dp5 = w;
    w->cap = c->max_batch;
    w->n = c->n;
    const size_t sz = (size_t)c->n * w->cap;
    for (int j = 0; j < 7; ++j) TG_TRY(dev_alloc(c, &w->k[j], sz));
    TG_TRY(dev_alloc(c, &w->ys, sz));
    TG_TRY(dev_alloc(c, &w->y5, sz));
    TG_TRY(dev_alloc(c, &w->err, sz));
    TG_TRY(dev_alloc(c, &w->y, sz));
    TG_TRY(dev_alloc(c, &w->h, (size_t)w->cap));
    TG_TRY(dev_alloc(c, &w->acc, (size_t)w->cap));
    TG_TRY(dev_alloc(c, &w->md, (size_t)w->cap));
    TG_TRY(dev_alloc(c, &w->norms, (size_t)w->cap));
    TG_TRY(dev_alloc(c, &w->npart, (size_t)w->cap * relerr_blocks(c)));
    TG_TRY(dev_alloc(c, &w->nsched, (size_t)w->cap));
    TG_TRY(dev_alloc(c, &w->flux, (size_t)7 * w->cap));
    TG_TRY(dev_alloc(c, &w->running, (size_t)1));
    TG_HIP(c, hipHostMalloc((void**)&w->running_host, sizeof(int32_t), hipHostMallocDefault));
    return TG_OK;
}

void dp5_destroy(tg_ctx* c) {
    Dp5Work* w = work_of(c);
    if (!w) return;
    for (int j = 0; j < 7; ++j) dev_free(&w->k[j]);
    dev_free(&w->ys);
    dev_free(&w->y5);
    dev_free(&w->err);
    dev_free(&w->y);
    dev_free(&w->h);
    dev_free(&w->acc);
    dev_free(&w->md);
    dev_free(&w->norms);
    dev_free(&w->npart);
    dev_free(&w->sched);
    dev_free(&w->nsched);
    dev_free(&w->flux);
    dev_free(&w->hlog);
    dev_free(&w->running);
    if (w->running_host) (void)hipHostFree(w->running_host);
    dev_free(&w->dense);
    delete w;
    c->dp5 = nullptr;
}

// operator evaluation for the whole batch with the per-member baths in
// d_bath (device); kernel launches only.  flux (device, one per member)
// receives the e-ph electron energy flux of each member.  On the batched
// path members whose step is 0 this attempt (finished: gate = the
// per-member step sizes) are skipped: the controller reads nothing of theirs.
static tg_status eval_batch(tg_ctx* c, const double* d_in, double* d_out, int32_t B,
                            const double* d_bath, int32_t flags, const double* gate, double* flux) {
    const int conserve = (flags & TG_CONSERVE) ? 1 : 0;
    if (flags & TG_PREC_MIXED) {
        TG_TRY(mixed_eval_device(c, d_in, d_out, B, conserve, gate, flux));
    } else {
        for (int b = 0; b < B; ++b)
            TG_TRY(plain_apply_device(c, d_in + (int64_t)b * c->n, d_out + (int64_t)b * c->n, 0.0, conserve,
                                      d_bath + b, flux + b));
        gate = nullptr;
    }
    return ee_apply_device(c, d_in, d_out, B, gate);
}

static tg_status finish_attempt(tg_ctx* c, Dp5Work* w, int32_t B, const Ctl& ctl, double* d_dense,
                                double* d_hlog);

// the device work of one attempted step (captured into a graph)
static tg_status attempt(tg_ctx* c, Dp5Work* w, int32_t B, const Ctl& ctl, double* d_bath,
                         int32_t flags, double rtol, double atol, double* d_dense, double* d_hlog) {
    hipStream_t st = c->stream;
    const int64_t n = c->n;
    double** k = w->k;
    const dim3 blk(256);
    const dim3 grd(ceil_div(n, 256 * 4) > 0 ? ceil_div(n, 256 * 4) : 1, B);
    // stage state y + h sum a_j k_j at the stage time t + cs h (the bath
    // schedule, if any, is evaluated there)
    auto stage = [&](std::initializer_list<double> coefs, int nk, double* out, double cs) -> tg_status {
        StageArgs a{};
        int j = 0;
        for (double v : coefs) a.a[j++] = v;
        for (int q = 0; q < nk; ++q) a.k[q] = k[q];
        a.nk = nk;
        hipLaunchKernelGGL(stage_kernel, grd, blk, 0, st, n, B, w->y, a, w->h, out, ctl.bath, w->md, cs,
                           d_bath);
        TG_CHECK_LAUNCH(c);
        return TG_OK;
    };
    auto eval = [&](const double* in, int s) {
        return eval_batch(c, in, k[s], B, d_bath, flags, w->h, w->flux + (int64_t)s * w->cap);
    };
    const dim3 cgrd(ceil_div(B, 256));
    hipLaunchKernelGGL(prep_kernel, cgrd, dim3(256), 0, st, w->md, ctl, w->h, w->running);
    TG_CHECK_LAUNCH(c);
    if ((flags & TG_PREC_MIXED) && mixed_dp5_fusable(c)) {
        // the same steps as below with the vector work between GEMMs fused
        // into one member pass each (Dp5PassArgs); no err vector
        TG_TRY(mixed_dp5_compact(c, w->h, B));
        Dp5PassArgs p{};
        p.y = w->y;
        p.h = w->h;
        p.md = w->md;
        p.bs = ctl.bath;
        p.bath = d_bath;
        p.conserve = (flags & TG_CONSERVE) ? 1 : 0;
        p.conn = c->d_conn;
        p.nel = c->nel;
        p.dim = c->dim;
        // finish k_s (index s-1 of k) evaluated at `fin`, then the next stage
        // state y + sum (a_j h) k_j over the k's in `terms`, k_s last
        // cfin: the coefficients of fin's own sum over the first of `terms`
        // (the previous pass's), so the pass recomputes fin instead of reading it
        auto pass = [&](int mode, int s_fin, const double* fin, std::initializer_list<int> terms,
                        std::initializer_list<double> ca, double* ys, double cs,
                        std::initializer_list<double> cfin = {}) -> tg_status {
            p.fin = fin;
            p.nfin = 0;
            for (double v : cfin) p.cfin[p.nfin++] = v;
            p.kout = s_fin > 0 ? k[s_fin - 1] : nullptr;
            p.flux = s_fin > 0 ? w->flux + (int64_t)(s_fin - 1) * w->cap : nullptr;
            int j = 0;
            for (int t : terms) p.kg[j++] = k[t];
            p.ng = j;
            j = 0;
            for (double v : ca) p.ca[j++] = v;
            p.ys = ys;
            p.cs = cs;
            TG_TRY(mixed_dp5_pass(c, mode, p, B));
            return mode == kDp5Last ? TG_OK : mixed_dp5_gemm(c);
        };
        // stages k2..k6 (DormandPrince54.hpp:565-574)
        TG_TRY(pass(kDp5Stage, 0, nullptr, {0}, {A21}, w->ys, C2));
        TG_TRY(pass(kDp5Mid, 2, w->ys, {0}, {A31, A32}, w->ys, C3, {A21}));
        TG_TRY(pass(kDp5Mid, 3, w->ys, {0, 1}, {A41, A42, A43}, w->ys, C4, {A31, A32}));
        TG_TRY(pass(kDp5Mid, 4, w->ys, {0, 1, 2}, {A51, A52, A53, A54}, w->ys, C5, {A41, A42, A43}));
        TG_TRY(pass(kDp5Mid, 5, w->ys, {0, 1, 2, 3}, {A61, A62, A63, A64, A65}, w->ys, 1.0,
                    {A51, A52, A53, A54}));
        // y5 (b2 = 0) and the y4 prefix, then k7 (:576-578), both at t + h
        const double cb[5] = {BH1, BH3, BH4, BH5, BH6};
        for (int q = 0; q < 5; ++q) p.cb[q] = cb[q];
        p.y4 = w->err;
        TG_TRY(pass(kDp5Mid, 6, w->ys, {0, 2, 3, 4}, {B1, B3, B4, B5, B6}, w->y5, 1.0));
        // k7, err = y5 - y4 (:580) and the error-norm sums
        p.c7 = BH7;
        p.rtol = rtol;
        p.atol = atol;
        p.norms = ctl.replay ? nullptr : w->norms;
        TG_TRY(pass(kDp5Last, 7, w->y5, {}, {}, nullptr, 1.0));
        return finish_attempt(c, w, B, ctl, d_dense, d_hlog);
    }
    // stages k2..k6 (DormandPrince54.hpp:565-574)
    TG_TRY(stage({A21}, 1, w->ys, C2));
    TG_TRY(eval(w->ys, 1));
    TG_TRY(stage({A31, A32}, 2, w->ys, C3));
    TG_TRY(eval(w->ys, 2));
    TG_TRY(stage({A41, A42, A43}, 3, w->ys, C4));
    TG_TRY(eval(w->ys, 3));
    TG_TRY(stage({A51, A52, A53, A54}, 4, w->ys, C5));
    TG_TRY(eval(w->ys, 4));
    TG_TRY(stage({A61, A62, A63, A64, A65}, 5, w->ys, 1.0));
    TG_TRY(eval(w->ys, 5));
    // y5 (b2 = 0) and k7 (:576-578), both at t + h
    TG_TRY(stage({B1, 0.0, B3, B4, B5, B6}, 6, w->y5, 1.0));
    TG_TRY(eval(w->y5, 6));
    // err = y5 - y4 (:580)
    {
        StageArgs a{};
        const double bh[7] = {BH1, 0.0, BH3, BH4, BH5, BH6, BH7};
        for (int q = 0; q < 7; ++q) {
            a.a[q] = bh[q];
            a.k[q] = k[q];
        }
        a.nk = 7;
        hipLaunchKernelGGL(error_kernel, grd, blk, 0, st, n, B, w->y, a, w->h, w->y5, w->err);
        TG_CHECK_LAUNCH(c);
    }
    if (!ctl.replay) TG_TRY(relerr_sums_device(c, w->err, w->y5, B, rtol, atol, w->npart, w->norms));
    return finish_attempt(c, w, B, ctl, d_dense, d_hlog);
}

// controller, dense outputs, accept and the running flag (after the norms)
static tg_status finish_attempt(tg_ctx* c, Dp5Work* w, int32_t B, const Ctl& ctl, double* d_dense,
                                double* d_hlog) {
    hipStream_t st = c->stream;
    const int64_t n = c->n;
    double** k = w->k;
    const dim3 blk(256);
    const dim3 grd(ceil_div(n, 256 * 4) > 0 ? ceil_div(n, 256 * 4) : 1, B);
    const dim3 cgrd(ceil_div(B, 256));
    hipLaunchKernelGGL(controller_kernel, cgrd, dim3(256), 0, st, w->md, ctl, w->norms, w->acc,
                       d_hlog, w->running);
    TG_CHECK_LAUNCH(c);
    if (d_dense) {
        hipLaunchKernelGGL(dense_kernel, dim3(ceil_div(n, 1024) > 0 ? ceil_div(n, 1024) : 1, B), blk,
                           0, st, n, w->md, ctl, w->y, k[0], k[2], k[3], k[4], k[5], k[6], d_dense);
        TG_CHECK_LAUNCH(c);
    }
    hipLaunchKernelGGL(accept_kernel, grd, blk, 0, st, n, w->acc, w->y, w->y5, k[0], k[6]);
    TG_CHECK_LAUNCH(c);
    TG_HIP(c, hipMemcpyAsync(w->running_host, w->running, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    return TG_OK;
}

}  // namespace tg

using namespace tg;

// Shared driver: adaptive when `sched` is null, otherwise member b takes the
// steps sched[b*h_cap .. + nsched[b]) and every attempt is accepted.
static tg_status dp5_run(tg_ctx* c, double* f, int64_t n, int32_t B, double t0, double t1,
                         double dense_dt, double rtol, double atol, const double* bath_n,
                         int32_t flags, double* dense_out, int32_t max_dense,
                         tg_step_stats* stats, double* h_acc, int32_t h_cap,
                         const double* sched, const int32_t* nsched) {
    if (!c || !f) return TG_ERR_ARG;
    if (n != c->n) return set_error(c, TG_ERR_ARG, "n mismatch");
    if (B < 1 || B > c->max_batch) return set_error(c, TG_ERR_ARG, "batch outside [1, max_batch]");
    if (!(dense_dt > 0.0)) return set_error(c, TG_ERR_ARG, "dense_dt must be > 0");
    if (!c->assembled) return set_error(c, TG_ERR_STATE, "tg_kernel_assemble not called");
    if (dense_out && max_dense < 1) return set_error(c, TG_ERR_ARG, "max_dense < 1");
    TG_TRY(dp5_ensure(c, B));
    Dp5Work* w = work_of(c);
    const bool dev = flags & TG_PTR_DEVICE;
    const bool quiet = flags & TG_QUIET;
    const bool replay = sched != nullptr;
    hipStream_t st = c->stream;
    const size_t vec = (size_t)n * B;

    // the bath: constant per member, or the schedule's n_b(t) (enforceValue at
    // t0 for k1 = timeOp(t0, enforce(t0, y0)), DormandPrince54.hpp:548)
    const bool sched_on = (flags & TG_BATH_SCHEDULE) != 0;
    if (sched_on && c->bs_ns < 2) return set_error(c, TG_ERR_STATE, "TG_BATH_SCHEDULE without tg_set_bath_schedule");
    if (sched_on && B > c->bs_batch)
        return set_error(c, TG_ERR_ARG, "batch %d exceeds the %d members of the bath schedule", B, c->bs_batch);
    BathSched bs{};
    if (sched_on) {
        bs.tab = c->h_bsched.data();   // host copy for the host-side calls below
        bs.ns = c->bs_ns;
        bs.t_start = c->bs_t0;
        bs.dt = c->bs_dt;
        bs.atol = c->bs_atol;
        bs.rtol = c->bs_rtol;
    }
    std::vector<double> bath(B);
    for (int b = 0; b < B; ++b)
        bath[b] = sched_on ? sched_value(bs.tab + (size_t)b * bs.ns, bs.ns, bs.t_start, bs.dt, t0)
                           : (bath_n ? bath_n[b] : c->bath_n);
    double* d_bath = c->d_bath;
    if (flags & TG_PREC_MIXED) {
        TG_TRY(mixed_set_bath(c, bath.data(), B));
        d_bath = mixed_bath_device(c);
    } else {
        TG_HIP(c, hipMemcpyAsync(c->d_bath, bath.data(), sizeof(double) * B, hipMemcpyHostToDevice, st));
    }

    TG_HIP(c, hipMemcpyAsync(w->y, f, sizeof(double) * vec,
                             dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
    double* d_dense = nullptr;
    if (dense_out) {
        if (dev) {
            d_dense = dense_out;
        } else {
            const size_t need = (size_t)max_dense * vec;
            if (w->dense_cap < need) {
                dev_free(&w->dense);
                TG_TRY(dev_alloc(c, &w->dense, need));
                w->dense_cap = need;
            }
            d_dense = w->dense;
        }
    }
    const int32_t log_cap = replay ? h_cap : (h_acc ? h_cap : 1);
    if (replay || h_acc) {
        const size_t need = (size_t)B * log_cap;
        if (replay && w->sched_cap < need) {
            dev_free(&w->sched);
            TG_TRY(dev_alloc(c, &w->sched, need));
            w->sched_cap = need;
        }
        if (h_acc && w->hlog_cap < need) {
            dev_free(&w->hlog);
            TG_TRY(dev_alloc(c, &w->hlog, need));
            w->hlog_cap = need;
        }
    }
    if (h_acc)  // entries past a member's accepted count keep the caller's values
        TG_HIP(c, hipMemcpyAsync(w->hlog, h_acc, sizeof(double) * (size_t)B * h_cap,
                                 hipMemcpyHostToDevice, st));
    if (replay) {
        TG_HIP(c, hipMemcpyAsync(w->sched, sched, sizeof(double) * B * h_cap, hipMemcpyHostToDevice, st));
        TG_HIP(c, hipMemcpyAsync(w->nsched, nsched, sizeof(int32_t) * B, hipMemcpyHostToDevice, st));
    }

    // controller state (DormandPrince54.hpp:527-543)
    Ctl ctl{};
    ctl.t1 = t1;
    ctl.dense_dt = dense_dt;
    ctl.min_h = dense_dt / 10000.;
    ctl.denom = (double)((c->dim + 1) * (c->dim + 2)) * (double)c->nel;
    ctl.B = B;
    ctl.replay = replay ? 1 : 0;
    ctl.h_cap = log_cap;
    ctl.max_dense = d_dense ? max_dense : 0;
    ctl.sched = replay ? w->sched : nullptr;
    ctl.nsched = w->nsched;
    ctl.bath = bs;
    ctl.bath.tab = sched_on ? c->d_bsched : nullptr;
    ctl.flux = w->flux;
    ctl.flux_stride = w->cap;
    std::vector<MemberDev> ms(B);
    for (int b = 0; b < B; ++b) {
        MemberDev& m = ms[b];
        std::memset(&m, 0, sizeof(m));
        m.t = t0;
        m.h = dense_dt;
        // the first step is vetted too (DormandPrince54.hpp:562-564)
        if (!replay)
            m.h = std::fmin(m.h, step_vetoer(bs, sched_on ? bs.tab + (size_t)b * bs.ns : nullptr, t0, m.h, dense_dt));
        m.dense_time = t0 + dense_dt;
        m.olderror = 1.0;
        m.st.evals = 1;
    }
    TG_HIP(c, hipMemcpyAsync(w->md, ms.data(), sizeof(MemberDev) * B, hipMemcpyHostToDevice, st));
    TG_TRY(eval_batch(c, w->y, w->k[0], B, d_bath, flags, nullptr, w->flux));  // k1 = timeOp(y0)
    TG_HIP(c, hipStreamSynchronize(st));

    // the attempt as a graph; replayed one attempt ahead of the flag read
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    TG_HIP(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const bool timing = c->time_kernel;
    c->time_kernel = false;   // no event records inside the graph
    tg_status cap = attempt(c, w, B, ctl, d_bath, flags, rtol, atol, d_dense, h_acc ? w->hlog : nullptr);
    c->time_kernel = timing;
    const hipError_t ec = hipStreamEndCapture(st, &graph);
    if (cap != TG_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return cap;
    }
    if (ec != hipSuccess)
        return set_error(c, TG_ERR_HIP, "graph capture of a DP5 attempt: %s", hipGetErrorString(ec));
    hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess)
        return set_error(c, TG_ERR_HIP, "graph instantiation: %s", hipGetErrorString(ei));

    tg_status rc = TG_OK;
    hipEvent_t ev[2] = {nullptr, nullptr};
    for (auto& e : ev)
        if (rc == TG_OK && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            rc = set_error(c, TG_ERR_HIP, "event creation");
    // a hard limit on attempts guards against a controller that never ends
    const int64_t max_attempts = replay ? (int64_t)h_cap + 2 : 100000000;
    for (int64_t a = 0; rc == TG_OK; ++a) {
        if (a >= max_attempts) {
            rc = set_error(c, TG_ERR_STATE, "more than %lld DP5 attempts", (long long)max_attempts);
            break;
        }
        if (hipGraphLaunch(exec, st) != hipSuccess || hipEventRecord(ev[a & 1], st) != hipSuccess) {
            rc = set_error(c, TG_ERR_HIP, "graph launch of a DP5 attempt");
            break;
        }
        if (a == 0) continue;
        if (hipEventSynchronize(ev[(a - 1) & 1]) != hipSuccess) {
            rc = set_error(c, TG_ERR_HIP, "DP5 attempt failed on the device");
            break;
        }
        if (*(volatile int32_t*)w->running_host == 0) break;
    }
    const hipError_t es = hipStreamSynchronize(st);
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipGraphExecDestroy(exec);
    if (rc != TG_OK) return rc;
    if (es != hipSuccess) return set_error(c, TG_ERR_HIP, "DP5 attempts: %s", hipGetErrorString(es));

    TG_HIP(c, hipMemcpyAsync(ms.data(), w->md, sizeof(MemberDev) * B, hipMemcpyDeviceToHost, st));
    TG_HIP(c, hipMemcpyAsync(f, w->y, sizeof(double) * vec,
                             dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
    if (dense_out && !dev)
        TG_HIP(c, hipMemcpyAsync(dense_out, d_dense, sizeof(double) * (size_t)max_dense * vec,
                                 hipMemcpyDeviceToHost, st));
    if (h_acc)
        TG_HIP(c, hipMemcpyAsync(h_acc, w->hlog, sizeof(double) * (size_t)B * h_cap,
                                 hipMemcpyDeviceToHost, st));
    TG_HIP(c, hipStreamSynchronize(st));
    if (sched_on && (flags & TG_PREC_MIXED)) mixed_bath_forget(c);   // the stages rewrote the device baths
    if (!quiet && B == 1) printf("Number of adaptive steps: %lld\n", (long long)ms[0].st.attempted);
    for (int b = 0; b < B; ++b) {
        ms[b].st.failed = ms[b].failed ? 1 : 0;
        ms[b].st.bath_energy = ms[b].ledger;
        if (stats) stats[b] = ms[b].st;
        if (ms[b].nonfinite) rc = TG_ERR_NONFINITE;
        else if (ms[b].failed && rc == TG_OK) rc = TG_ERR_STEP_UNDERFLOW;
    }
    if (rc != TG_OK)
        set_error(c, rc, rc == TG_ERR_NONFINITE ? "non-finite error norm" : "step size underflow");
    return rc;
}

extern "C" tg_status tg_dp5_integrate(tg_ctx* c, double* f, int64_t n, int32_t B, double t0,
                                      double t1, double dense_dt, double rtol, double atol,
                                      const double* bath_n, int32_t flags, double* dense_out,
                                      int32_t max_dense, tg_step_stats* stats, double* h_accepted,
                                      int32_t h_cap) {
    if (h_accepted && h_cap < 1) return TG_ERR_ARG;
    return dp5_run(c, f, n, B, t0, t1, dense_dt, rtol, atol, bath_n, flags, dense_out, max_dense,
                   stats, h_accepted, h_cap, nullptr, nullptr);
}

extern "C" tg_status tg_dp5_replay(tg_ctx* c, double* f, int64_t n, int32_t B, double t0,
                                   double t1, double dense_dt, const double* h_steps,
                                   const int32_t* n_steps, int32_t h_cap, const double* bath_n,
                                   int32_t flags, double* dense_out, int32_t max_dense) {
    if (!c || !h_steps || !n_steps || h_cap < 1) return TG_ERR_ARG;
    for (int b = 0; b < B; ++b)
        if (n_steps[b] < 0 || n_steps[b] > h_cap)
            return set_error(c, TG_ERR_ARG, "n_steps[%d] outside [0, h_cap]", b);
    return dp5_run(c, f, n, B, t0, t1, dense_dt, 1.0, 1.0, bath_n, flags | TG_QUIET, dense_out,
                   max_dense, nullptr, nullptr, h_cap, h_steps, n_steps);
}
