// tg_dp5.hip -- adaptive Dormand-Prince 5(4) over a batch of members.
//
// Restates Algorithms::dormandPrince54 (vetoer form,
// include/Generics/Algorithms/DormandPrince54.hpp:479-665) for B independent
// members that share the collision operator: every member keeps its own time,
// step size, PI-controller memory and reject flag; an attempted step runs the
// six stage evaluations for all unfinished members at once (one batched
// operator call per stage), then each member is accepted or rejected on its
// own error norm (Band::relativeError, Band.cpp:87-101).  enforceValue and the
// step vetoer are the identity (no constrained bands in this operator).
//
// Device work per attempted step: 6 stage-sum kernels fused with nothing else
// to read (y, k1..k6 -> ys), 6 operator evaluations, one fused y4/err kernel,
// one batched error-norm reduction; after the host decision one dense-output
// kernel (all dense points of all members) and one accept kernel.
#include <cmath>
#include <cstring>
#include <vector>

#include "tg_internal.hpp"

namespace tg {

// Butcher tableau (DormandPrince54.hpp:503-511)
constexpr double C2 = 1. / 5., C3 = 3. / 10., C4 = 4. / 5., C5 = 8. / 9.;
constexpr double A21 = 1. / 5.;
constexpr double A31 = 3. / 40., A32 = 9. / 40.;
constexpr double A41 = 44. / 45., A42 = -56. / 15., A43 = 32. / 9.;
constexpr double A51 = 19372. / 6561., A52 = -25360. / 2187., A53 = 64448. / 6561., A54 = -212. / 729.;
constexpr double A61 = 9017. / 3168., A62 = -355. / 33., A63 = 46732. / 5247., A64 = 49. / 176.,
                 A65 = -5103. / 18656.;
constexpr double B1 = 35. / 384., B3 = 500. / 1113., B4 = 125. / 192., B5 = -2187. / 6784., B6 = 11. / 84.;
constexpr double BH1 = 5179. / 57600., BH3 = 7571. / 16695., BH4 = 393. / 640.,
                 BH5 = -92097. / 339200., BH6 = 187. / 2100., BH7 = 1. / 40.;

struct StageArgs {
    const double* k[7];   // stage derivatives (n x B each)
    double a[7];          // tableau coefficients (multiplied by h per member)
    int nk;
};

// out[:, b] = y[:, b] + sum_j (a_j h_b) k_j[:, b]; the product a_j*h_b is
// formed first, as the reference does ((a_ij * h) * k_j).
__global__ void stage_kernel(int64_t n, int B, const double* __restrict__ y, StageArgs s,
                             const double* __restrict__ h, double* __restrict__ out) {
    const int b = blockIdx.y;
    const double hb = h[b];
    double c[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) c[j] = s.a[j] * hb;
    const int64_t base = (int64_t)b * n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double v = y[base + i];
        for (int j = 0; j < s.nk; ++j) v += c[j] * s.k[j][base + i];
        out[base + i] = v;
    }
}

// err = y5 - (y + sum_j (h_b bh_j) k_j), j over k1,k3..k7
__global__ void error_kernel(int64_t n, int B, const double* __restrict__ y, StageArgs s,
                             const double* __restrict__ h, const double* __restrict__ y5,
                             double* __restrict__ err) {
    const int b = blockIdx.y;
    const double hb = h[b];
    double c[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) c[j] = hb * s.a[j];
    const int64_t base = (int64_t)b * n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double v = y[base + i];
        for (int j = 0; j < s.nk; ++j) v += c[j] * s.k[j][base + i];
        err[base + i] = y5[base + i] - v;
    }
}

// One dense output point: member, six coefficients (s*h*btilde_j), slot.
struct DensePoint {
    int32_t member;
    int32_t pad;
    int64_t dst;          // element offset into the dense buffer
    double c[6];          // for k1, k3, k4, k5, k6, k7
};

__global__ void dense_kernel(int64_t n, const double* __restrict__ y, const double* __restrict__ k1,
                             const double* __restrict__ k3, const double* __restrict__ k4,
                             const double* __restrict__ k5, const double* __restrict__ k6,
                             const double* __restrict__ k7, const DensePoint* __restrict__ pts,
                             double* __restrict__ out) {
    const DensePoint p = pts[blockIdx.y];
    const int64_t base = (int64_t)p.member * n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = base + i;
        out[p.dst + i] = y[q] + p.c[0] * k1[q] + p.c[1] * k3[q] + p.c[2] * k4[q] +
                         p.c[3] * k5[q] + p.c[4] * k6[q] + p.c[5] * k7[q];
    }
}

// accepted members: y <- y5, k1 <- k7 (FSAL)
__global__ void accept_kernel(int64_t n, const int8_t* __restrict__ acc, double* __restrict__ y,
                              const double* __restrict__ y5, double* __restrict__ k1,
                              const double* __restrict__ k7) {
    const int b = blockIdx.y;
    if (!acc[b]) return;
    const int64_t base = (int64_t)b * n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        y[base + i] = y5[base + i];
        k1[base + i] = k7[base + i];
    }
}

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
    DensePoint* pts = nullptr;
    int32_t pts_cap = 0;
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
    dev_free(&w->pts);
    dev_free(&w->dense);
    delete w;
    c->dp5 = nullptr;
}

// operator evaluation for the whole batch
static tg_status eval_batch(tg_ctx* c, const double* d_in, double* d_out, int32_t B,
                            const std::vector<double>& bath, int32_t flags) {
    const int conserve = (flags & TG_CONSERVE) ? 1 : 0;
    if (flags & TG_PREC_MIXED) return mixed_apply_device(c, d_in, d_out, B, bath.data(), conserve);
    for (int b = 0; b < B; ++b)
        TG_TRY(plain_apply_device(c, d_in + (int64_t)b * c->n, d_out + (int64_t)b * c->n, bath[b],
                                  conserve));
    return TG_OK;
}

struct MemberState {
    double t, h, dense_time, olderror;
    bool rejected, failed, nonfinite;
    tg_step_stats st;
};

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
    hipStream_t st = c->stream;
    const size_t vec = (size_t)n * B;

    std::vector<double> bath(B);
    for (int b = 0; b < B; ++b) bath[b] = bath_n ? bath_n[b] : c->bath_n;

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

    // controller state (DormandPrince54.hpp:527-543)
    const double min_h = dense_dt / 10000.;
    const double minscale = 0.333333, maxscale = 3.0, safety = 0.98;
    const double alpha = 0.7 / 5., beta = 0.4 / 5.;
    std::vector<MemberState> ms(B);
    for (auto& m : ms) {
        m.t = t0;
        m.h = dense_dt;
        m.dense_time = t0 + dense_dt;
        m.olderror = 1.0;
        m.rejected = m.failed = m.nonfinite = false;
        std::memset(&m.st, 0, sizeof(m.st));
        m.st.evals = 1;
    }

    double** k = w->k;
    TG_TRY(eval_batch(c, w->y, k[0], B, bath, flags));  // k1 = timeOp(y0)

    std::vector<double> hh(B), en(B);
    std::vector<int8_t> acc(B);
    std::vector<DensePoint> pts;
    const dim3 blk(256);
    const dim3 grd(ceil_div(n, 256 * 4) > 0 ? ceil_div(n, 256 * 4) : 1, B);
    auto stage = [&](const double* in_y, std::initializer_list<double> coefs, int nk,
                     double* out) -> tg_status {
        StageArgs a{};
        int j = 0;
        for (double v : coefs) a.a[j++] = v;
        for (int q = 0; q < nk; ++q) a.k[q] = k[q];
        a.nk = nk;
        hipLaunchKernelGGL(stage_kernel, grd, blk, 0, st, n, B, in_y, a, w->h, out);
        TG_CHECK_LAUNCH(c);
        return TG_OK;
    };

    const bool replay = sched != nullptr;
    std::vector<int32_t> step_idx(B, 0);
    auto is_active = [&](int b) {
        const MemberState& m = ms[b];
        if (replay) return step_idx[b] < nsched[b] && m.t < t1;
        return m.t < t1 && !m.failed;
    };
    while (true) {
        bool any = false;
        for (int b = 0; b < B; ++b) {
            MemberState& m = ms[b];
            const bool active = is_active(b);
            if (active && replay) m.h = sched[(int64_t)b * h_cap + step_idx[b]];
            hh[b] = active ? m.h : 0.0;
            any |= active;
        }
        if (!any) break;
        TG_HIP(c, hipMemcpyAsync(w->h, hh.data(), sizeof(double) * B, hipMemcpyHostToDevice, st));
        // stages k2..k6 (DormandPrince54.hpp:565-574)
        TG_TRY(stage(w->y, {A21}, 1, w->ys));
        TG_TRY(eval_batch(c, w->ys, k[1], B, bath, flags));
        TG_TRY(stage(w->y, {A31, A32}, 2, w->ys));
        TG_TRY(eval_batch(c, w->ys, k[2], B, bath, flags));
        TG_TRY(stage(w->y, {A41, A42, A43}, 3, w->ys));
        TG_TRY(eval_batch(c, w->ys, k[3], B, bath, flags));
        TG_TRY(stage(w->y, {A51, A52, A53, A54}, 4, w->ys));
        TG_TRY(eval_batch(c, w->ys, k[4], B, bath, flags));
        TG_TRY(stage(w->y, {A61, A62, A63, A64, A65}, 5, w->ys));
        TG_TRY(eval_batch(c, w->ys, k[5], B, bath, flags));
        // y5 (b2 = 0) and k7 (:576-578)
        TG_TRY(stage(w->y, {B1, 0.0, B3, B4, B5, B6}, 6, w->y5));
        TG_TRY(eval_batch(c, w->y5, k[6], B, bath, flags));
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
        if (replay) std::fill(en.begin(), en.end(), 0.0);
        else TG_TRY(relative_error_device(c, w->err, w->y5, B, rtol, atol, en.data()));

        pts.clear();
        for (int b = 0; b < B; ++b) {
            MemberState& m = ms[b];
            acc[b] = 0;
            if (!is_active(b)) continue;
            if (replay) ++step_idx[b];
            m.st.attempted += 1;
            m.st.evals += 6;
            const double e = en[b];
            const double h = m.h;
            if (!std::isfinite(e)) {
                m.failed = true;
                m.nonfinite = true;
                continue;
            }
            if (!quiet && B == 1) printf("** Time: %f timestep: %f\n", m.t, h);
            if (e <= 1.0) {
                // dense output (:598-611)
                while (m.dense_time <= m.t + 1.0001 * h && m.dense_time <= t1 + .0001 * h) {
                    const double s = (m.dense_time - m.t) / h;
                    if (d_dense && m.st.n_dense < max_dense) {
                        DensePoint p{};
                        p.member = b;
                        p.dst = ((int64_t)b * max_dense + m.st.n_dense) * n;
                        const double bt1 = -(435. * s * s * s - 1184. * s * s + 1098. * s - 384.) / (384.);
                        const double bt3 = 500. * s * (6. * s * s - 14. * s + 9.) / (1113.);
                        const double bt4 = -125. * s * (9. * s * s - 16. * s + 6.) / (192.);
                        const double bt5 = 729. * s * (35. * s * s - 64. * s + 26.) / (6784.);
                        const double bt6 = -11. * s * (3. * s - 2.) * (5. * s - 6.) / (84.);
                        const double bt7 = s * (s - 1.) * (5. * s - 3.) / (2.);
                        p.c[0] = s * h * bt1;
                        p.c[1] = s * h * bt3;
                        p.c[2] = s * h * bt4;
                        p.c[3] = s * h * bt5;
                        p.c[4] = s * h * bt6;
                        p.c[5] = s * h * bt7;
                        pts.push_back(p);
                    }
                    m.st.n_dense += 1;
                    m.dense_time += dense_dt;
                }
                if (h_acc && m.st.accepted < h_cap) h_acc[(int64_t)b * h_cap + m.st.accepted] = h;
                m.t += h;
                acc[b] = 1;
                double scale;
                if (e == 0.) {
                    scale = maxscale;
                } else {
                    scale = safety * std::pow(1. / e, alpha) * std::pow(m.olderror, beta);
                    scale = std::fmin(std::fmax(scale, minscale), maxscale);
                }
                m.h = m.rejected ? h * std::fmin(1.0, scale) : h * scale;
                m.rejected = false;
                m.olderror = e;
                m.st.accepted += 1;
            } else {
                double scale = safety * std::pow(1. / e, 1 / 5.);
                scale = std::fmin(std::fmax(scale, minscale), maxscale);
                m.h = h * scale;
                m.rejected = true;
                m.st.rejected += 1;
                if (m.h < min_h) m.failed = true;
            }
        }
        if (!pts.empty()) {
            if ((int32_t)pts.size() > w->pts_cap) {
                dev_free(&w->pts);
                w->pts_cap = (int32_t)pts.size() * 2;
                TG_TRY(dev_alloc(c, &w->pts, (size_t)w->pts_cap));
            }
            TG_HIP(c, hipMemcpyAsync(w->pts, pts.data(), sizeof(DensePoint) * pts.size(),
                                     hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(dense_kernel, dim3(ceil_div(n, 1024) > 0 ? ceil_div(n, 1024) : 1, pts.size()),
                               blk, 0, st, n, w->y, k[0], k[2], k[3], k[4], k[5], k[6], w->pts, d_dense);
            TG_CHECK_LAUNCH(c);
            // pts is reused next step: wait for the copy before the host edits it
            TG_HIP(c, hipStreamSynchronize(st));
        }
        TG_HIP(c, hipMemcpyAsync(w->acc, acc.data(), B, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(accept_kernel, grd, blk, 0, st, n, w->acc, w->y, w->y5, k[0], k[6]);
        TG_CHECK_LAUNCH(c);
        TG_HIP(c, hipStreamSynchronize(st));  // acc is host memory reused next step
    }
    if (!quiet && B == 1) printf("Number of adaptive steps: %lld\n", (long long)ms[0].st.attempted);

    TG_HIP(c, hipMemcpyAsync(f, w->y, sizeof(double) * vec,
                             dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
    if (dense_out && !dev)
        TG_HIP(c, hipMemcpyAsync(dense_out, d_dense, sizeof(double) * (size_t)max_dense * vec,
                                 hipMemcpyDeviceToHost, st));
    TG_HIP(c, hipStreamSynchronize(st));
    tg_status rc = TG_OK;
    for (int b = 0; b < B; ++b) {
        ms[b].st.failed = ms[b].failed ? 1 : 0;
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
                   max_dense, nullptr, nullptr, 1, h_steps, n_steps);
}
