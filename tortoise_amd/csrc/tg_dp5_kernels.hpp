// tg_dp5_kernels.hpp -- device side of the batched Dormand-Prince 5(4)
// driver, shared by the W-form driver (tg_dp5.hip) and the tuple operator's
// (tg_tuple.hip): tableau, per-member controller state and the stage, error,
// controller, dense-output and accept kernels.  The state of a member is a
// flat vector of n doubles; member b at b*n.
//
// Restates Algorithms::dormandPrince54 (vetoer form,
// include/Generics/Algorithms/DormandPrince54.hpp:479-665).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/tortoise_gpu.h"

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

// Time-dependent bath of a constrained phonon band: member b's occupation
// n_b(t) is given by samples at t_start + k dt, piecewise linear between them
// and held at the end values outside.  This is the device form of
// Band::constrainedPopulation (include/Physics/Material/Band.hpp:70-74) for a
// uniform population, which MaterialStatus::applyConstraints
// (src/Physics/Status/MaterialStatus.cpp:198-206) writes into the state at
// every stage time.
struct BathSched {
    const double* tab;    // [B][ns] samples (device memory in kernels); null: constant bath
    int32_t ns;           // samples per member (>= 2)
    double t_start, dt;   // time of sample 0, sample spacing
    double atol, rtol;    // the constrained band's absolute / relative error tolerances
};

// n_b(t) from member b's samples s[0..ns); the same operations on host and
// device (no contraction: the checker restates them operation by operation)
__host__ __device__ inline double sched_value(const double* s, int ns, double t_start, double dt,
                                              double t) {
#pragma clang fp contract(off)
    const double u = (t - t_start) / dt;
    if (!(u > 0.0)) return s[0];
    if (u >= (double)(ns - 1)) return s[ns - 1];
    int k = (int)floor(u);
    if (k > ns - 2) k = ns - 2;
    const double w = u - (double)k;
    return fma(w, s[k + 1] - s[k], s[k]);
}

// MaterialTimeStatus::timeStepVetoer (src/Physics/Status/MaterialTimeStatus.cpp:60-100)
// for one member, called with timesToTest = {t, t + c2 h, t + c3 h, t + c4 h,
// t + c5 h, t + h} as DormandPrince54.hpp:563-564,640-642 build it.  The
// constrained population is sampled at t + dense_dt, t + 2 dense_dt, ...;
// the first sample inside interval i that leaves the interval average by
// more than 1.3 |L - R| + atol + rtol |average| (L, R: the population at the
// interval's ends) caps the step at that sample's time.  Without a
// constrained band (s null) the cap is timesToTest.back() - timesToTest[0],
// as in the reference.  The population is uniform, so the max over nodes of
// the reference's nodal Function arithmetic is this scalar expression.
__host__ __device__ inline double step_vetoer(const BathSched& bs, const double* s, double t, double h,
                                              double dense_dt) {
#pragma clang fp contract(off)
    const double times[6] = {t, t + C2 * h, t + C3 * h, t + C4 * h, t + C5 * h, t + h};
    double crit = times[5];
    if (s) {
        const double var = 1. + 0.3;
        double cur = times[0] + dense_dt;
        double L = sched_value(s, bs.ns, bs.t_start, bs.dt, times[0]);
        double R = sched_value(s, bs.ns, bs.t_start, bs.dt, times[1]);
        bool found = false;
        for (int i = 0; i < 5 && !found; ++i) {
            if (i != 0) {
                L = R;
                R = sched_value(s, bs.ns, bs.t_start, bs.dt, times[i + 1]);
            }
            const double avg = 0.5 * (L + R);
            const double acc = var * fabs(L - R) + bs.atol + bs.rtol * fabs(avg);
            while (cur < times[i + 1]) {
                const double v = sched_value(s, bs.ns, bs.t_start, bs.dt, cur);
                if (fabs(v - avg) / acc > 1.) {
                    crit = cur;
                    found = true;
                    break;
                }
                const double next = cur + dense_dt;
                if (!(next > cur)) break;   // dense_dt below the resolution of t: no further samples
                cur = next;
            }
        }
    }
    return crit - times[0];
}

// Per-member controller state (DormandPrince54.hpp:527-543, 547-661), on the
// device.  The accepted step's start, size and first dense time are kept for
// the dense-output kernel.  ledger: energy the phonon bath has received (the
// e-ph channel's electron energy flux, integrated with the step weights).
struct MemberDev {
    double t, h, dense_time, olderror;
    double step_t, step_h, step_dense_time;
    double ledger;
    int32_t rejected, failed, nonfinite, active;
    int32_t step_idx, acc, dense_k0, dense_k1;
    tg_step_stats st;
};

// One fused member pass of the batched DP5 on the mixed-precision path
// (tg_mixed.hip, mixed_dp5_pass): the vector work between two GEMMs of an
// attempted step for every member at once.
//   kDp5Stage  first stage of the attempt: the stage state ys = y + sum_j
//              (a_j h) k_j (ng terms) is formed in LDS, written out, packed
//   kDp5Mid    the stage derivative k_s = I(fin) is finished (epilogue,
//              number-defect removal, energy flux) and written out; the next
//              stage state y + sum_j (a_j h) k_j + (a_s h) k_s is formed with
//              k_s still in registers, written out and packed; before y5 the
//              pass also writes the y4 prefix y + sum_j (bh_j h) k_j + (bh6 h) k6
//   kDp5Last   k7 = I(y5) is finished and written out; err = y5 - (y4 prefix
//              + (bh7 h) k7) and r = err / (rtol min(|y5|, |1 - y5|) + atol)
//              go to LDS, and the member's error-norm sum (Band.cpp:87-101)
//              is taken over the elements
// Sums run left to right in ascending j with every product rounded
// separately, as DormandPrince54.hpp:565-580 evaluates them; terms whose
// coefficient is zero (a72 = b2 = bh2 = 0) are left out.
enum { kDp5Stage = 0, kDp5Mid = 1, kDp5Last = 2 };

struct Dp5PassArgs {
    const double* y;          // the attempt's start state (n x B)
    const double* kg[6];      // stage derivatives of the next stage's sum read from memory, ascending j
    double ca[7];             // their coefficients, then the finished k's (kDp5Mid)
    double cb[7];             // y4-prefix coefficients, same order (y4 written)
    int32_t ng;               // entries of kg
    double* ys;               // next stage state (kDp5Stage, kDp5Mid)
    double* y4;               // y4 prefix: written by the kDp5Mid pass before y5 (else null), read by kDp5Last
    double c7;                // kDp5Last: bh7
    const double* fin;        // kDp5Mid, kDp5Last: the state the finished k is evaluated at
    int32_t nfin;             // kDp5Mid: > 0 when fin = y + sum_{j < nfin} (cfin_j h) kg[j] with the
    double cfin[6];           //   terms this pass reads (the previous pass's sum, recomputed in the same
                              //   order instead of reading fin again for the number-defect weights)
    double* kout;             // the finished k
    double* flux;             // its e-ph electron energy flux per member (null: none)
    const double* h;          // per-member step of the attempt; 0: finished member, skipped
    const MemberDev* md;      // member times (bath schedule)
    BathSched bs;             // bath schedule (tab null: constant bath)
    double cs;                // next stage's time fraction: its bath is n_b(t + cs h)
    double* bath;             // per-member bath: read for the finished k, rewritten for the next stage
    double rtol, atol;        // kDp5Last: error-norm tolerances
    double* norms;            // kDp5Last: error-norm sum per member (null: no norm, replay)
    const int32_t* conn;      // mesh elements, nel x (dim + 1)
    int64_t nel;
    int32_t dim;
    int32_t conserve;         // number-defect removal
};

namespace {

// out[:, b] = y[:, b] + sum_j (a_j h_b) k_j[:, b]; the product a_j*h_b is
// formed first, as the reference does ((a_ij * h) * k_j), and every product
// and sum is rounded separately, left to right, as the reference's vector
// expression is (no fused multiply-adds: y5 - y4 is a difference of nearly
// equal vectors, and the error norm divides it by tolerances far below the
// resolution of filled states, so the rounding pattern is part of what the
// step controller sees; the stage, error and dense kernels round exactly
// like DormandPrince54.hpp:565-608).  With a bath
// schedule, block 0 of each member also writes the member's bath occupation
// at the stage time t + c h (enforceValue at the stage time,
// DormandPrince54.hpp:565-578) for the operator evaluation that follows.
__global__ void stage_kernel(int64_t n, int B, const double* __restrict__ y, StageArgs s,
                             const double* __restrict__ h, double* __restrict__ out,
                             BathSched bs = BathSched{}, const MemberDev* __restrict__ md = nullptr,
                             double cs = 0.0, double* __restrict__ bath = nullptr) {
#pragma clang fp contract(off)
    const int b = blockIdx.y;
    const double hb = h[b];
    if (hb == 0.0) return;   // finished member: nothing of this attempt is read
    if (bs.tab && blockIdx.x == 0 && threadIdx.x == 0) {
        const double ts = md[b].t + cs * hb;
        bath[b] = sched_value(bs.tab + (int64_t)b * bs.ns, bs.ns, bs.t_start, bs.dt, ts);
    }
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
#pragma clang fp contract(off)
    const int b = blockIdx.y;
    const double hb = h[b];
    if (hb == 0.0) return;   // finished member: the controller skips it
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

struct Ctl {
    double t1, dense_dt, min_h, denom;
    int32_t B, replay, h_cap, max_dense;
    const double* sched;      // replay: step sizes, B x h_cap
    const int32_t* nsched;    // replay: steps per member
    BathSched bath;           // time-dependent bath (tab null: constant)
    double* flux;             // null, or per stage s = 0..6 (k1..k7) and member the e-ph
    int64_t flux_stride;      //   electron energy flux at flux[s * flux_stride + b]
};

__device__ inline bool member_running(const MemberDev& m, const Ctl& k, int b) {
    if (k.replay) return m.step_idx < k.nsched[b] && m.t < k.t1;
    return m.t < k.t1 && !m.failed;
}

// members still running take their step size, the others 0; clears the
// "any member running" flag the controller sets
__global__ void prep_kernel(MemberDev* __restrict__ md, Ctl k, double* __restrict__ h,
                            int32_t* __restrict__ running) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b == 0) *running = 0;
    if (b >= k.B) return;
    MemberDev m = md[b];
    m.active = member_running(m, k, b) ? 1 : 0;
    if (m.active && k.replay) m.h = k.sched[(int64_t)b * k.h_cap + m.step_idx];
    h[b] = m.active ? m.h : 0.0;
    md[b] = m;
}

// accept / reject per member (DormandPrince54.hpp:582-661); norms holds the
// error-norm sums, err = sqrt(sum / denom) (Band.cpp:87-101)
__global__ void controller_kernel(MemberDev* __restrict__ md, Ctl k,
                                  const double* __restrict__ norms, int8_t* __restrict__ acc,
                                  double* __restrict__ hlog, int32_t* __restrict__ running) {
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < k.B) {
        const double minscale = 0.333333, maxscale = 3.0, safety = 0.98;
        const double alpha = 0.7 / 5., beta = 0.4 / 5.;
        MemberDev m = md[b];
        m.acc = 0;
        m.dense_k0 = m.dense_k1 = m.st.n_dense;
        if (m.active) {
            if (k.replay) ++m.step_idx;
            m.st.attempted += 1;
            m.st.evals += 6;
            const double e = k.replay ? 0.0 : sqrt(norms[b] / k.denom);
            const double h = m.h;
            if (!isfinite(e)) {
                // a trial step so far outside the stability region that its
                // stages overflow: rejected with the smallest scale (what
                // DormandPrince54.hpp:646-649 gives for an infinite norm); a
                // member that cannot recover fails by step underflow, reported
                // as a non-finite error norm
                m.h = h * minscale;
                m.rejected = 1;
                m.st.rejected += 1;
                if (m.h < k.min_h) {
                    m.failed = 1;
                    m.nonfinite = 1;
                }
            } else if (e <= 1.0) {
                // dense outputs inside the step (:598-611)
                m.step_t = m.t;
                m.step_h = h;
                m.step_dense_time = m.dense_time;
                while (m.dense_time <= m.t + 1.0001 * h && m.dense_time <= k.t1 + .0001 * h) {
                    m.st.n_dense += 1;
                    m.dense_time += k.dense_dt;
                }
                m.dense_k1 = m.st.n_dense;
                if (hlog && m.st.accepted < k.h_cap) hlog[(int64_t)b * k.h_cap + m.st.accepted] = h;
                m.t += h;
                m.acc = 1;
                double scale;
                if (e == 0.) {
                    scale = maxscale;
                } else {
                    scale = safety * pow(1. / e, alpha) * pow(m.olderror, beta);
                    scale = fmin(fmax(scale, minscale), maxscale);
                }
                m.h = m.rejected ? h * fmin(1.0, scale) : h * scale;
                // the vetoer caps the next step (:640-642)
                if (!k.replay)
                    m.h = fmin(m.h, step_vetoer(k.bath, k.bath.tab ? k.bath.tab + (int64_t)b * k.bath.ns : nullptr,
                                                m.t, m.h, k.dense_dt));
                m.rejected = 0;
                m.olderror = e;
                m.st.accepted += 1;
                if (k.flux) {
                    // bath ledger: minus the electron energy the step's weights
                    // integrate (y5 = y + sum_j (h b_j) k_j), then FSAL k1 <- k7
                    double* ef = k.flux;
                    const int64_t S = k.flux_stride;
                    const double de = (h * B1) * ef[b] + (h * B3) * ef[2 * S + b] + (h * B4) * ef[3 * S + b] +
                                      (h * B5) * ef[4 * S + b] + (h * B6) * ef[5 * S + b];
                    m.ledger -= de;
                    ef[b] = ef[6 * S + b];
                }
            } else {
                double scale = safety * pow(1. / e, 1 / 5.);
                scale = fmin(fmax(scale, minscale), maxscale);
                m.h = h * scale;
                m.rejected = 1;
                m.st.rejected += 1;
                if (m.h < k.min_h) m.failed = 1;
            }
        }
        acc[b] = (int8_t)m.acc;
        if (member_running(m, k, b)) atomicOr(&any, 1);
        md[b] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0 && any) atomicOr(running, 1);
}

// dense outputs of the step just accepted: y(t_s) = y + sum_j (s h bt_j(s)) k_j
// with s = (t_s - t)/h (DormandPrince54.hpp:601-608)
__global__ void dense_kernel(int64_t n, const MemberDev* __restrict__ md, Ctl k,
                             const double* __restrict__ y, const double* __restrict__ k1,
                             const double* __restrict__ k3, const double* __restrict__ k4,
                             const double* __restrict__ k5, const double* __restrict__ k6,
                             const double* __restrict__ k7, double* __restrict__ out) {
#pragma clang fp contract(off)
    const int b = blockIdx.y;
    const MemberDev& m = md[b];
    if (!m.acc || m.dense_k0 >= m.dense_k1 || m.dense_k0 >= k.max_dense) return;
    const int64_t base = (int64_t)b * n;
    const double t = m.step_t, h = m.step_h;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = base + i;
        double td = m.step_dense_time;
        for (int32_t slot = m.dense_k0; slot < m.dense_k1 && slot < k.max_dense; ++slot) {
            const double s = (td - t) / h;
            const double bt1 = -(435. * s * s * s - 1184. * s * s + 1098. * s - 384.) / (384.);
            const double bt3 = 500. * s * (6. * s * s - 14. * s + 9.) / (1113.);
            const double bt4 = -125. * s * (9. * s * s - 16. * s + 6.) / (192.);
            const double bt5 = 729. * s * (35. * s * s - 64. * s + 26.) / (6784.);
            const double bt6 = -11. * s * (3. * s - 2.) * (5. * s - 6.) / (84.);
            const double bt7 = s * (s - 1.) * (5. * s - 3.) / (2.);
            out[((int64_t)b * k.max_dense + slot) * n + i] =
                y[q] + (s * h * bt1) * k1[q] + (s * h * bt3) * k3[q] + (s * h * bt4) * k4[q] +
                (s * h * bt5) * k5[q] + (s * h * bt6) * k6[q] + (s * h * bt7) * k7[q];
            td += k.dense_dt;
        }
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

}  // namespace
}  // namespace tg
