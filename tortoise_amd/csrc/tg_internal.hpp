// tg_internal.hpp -- context layout and helpers shared by the HIP sources of
// libtortoise_gpu.so.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/tortoise_gpu.h"
#include "tg_hat.hpp"
#include "tg_kentry.hpp"

namespace tg {

struct Buffer {
    void* p = nullptr;
    size_t bytes = 0;
};

struct tg_ctx_impl;

}  // namespace tg

struct tg_ctx {
    // ---- problem
    int dim = 2;
    int32_t nsec[3] = {1, 1, 1};
    double origin[3] = {0, 0, 0};
    double extent[3] = {1, 1, 1};
    int64_t n = 0;       // nodes
    int64_t nel = 0;     // elements
    double vol = 0;      // element volume
    double g = 0, omega0 = 0, sigma = 0, bath_n = 0, de = 0, cutoff = 8;
    int32_t max_batch = 1;
    // occupation factors of the test node: TG_OCC_NODAL or TG_OCC_HAT; with
    // TG_OCC_HAT d_hat holds the packed hat stencil (tg_hat.hpp)
    int32_t occ = TG_OCC_NODAL;
    uint32_t* d_hat = nullptr;
    double* d_xb = nullptr;        // scratch: one state's occupation factors (caller order)

    // energy grid and projection
    double e0 = 0;
    int32_t nbins = 0;
    int32_t maxw = 0;     // projection window width
    int32_t m_lo = 0;     // first sampled offset of the broadened delta
    int32_t nm = 0;       // number of sampled offsets
    int32_t qw = 0;       // convolved window width = maxw + nm - 1

    // ---- device data
    double* d_eps = nullptr;
    double* d_V = nullptr;
    int32_t* d_conn = nullptr;     // nel x (dim+1)
    int32_t* d_off = nullptr;      // projection window start per node
    int32_t* d_len = nullptr;      // projection window length per node
    double* d_R = nullptr;         // n x maxw
    double* d_Q = nullptr;         // n x qw (kept: K's entries are computed from R and Q where
                                   // needed, K itself is never stored densely)
    double* d_gm = nullptr;        // nm
    double* d_rk = nullptr;        // K 1
    double* d_ck = nullptr;        // K^T 1
    int64_t nnz = 0;               // structurally nonzero entries of K
    // exact nonzero extents of K in energy order (-1 when empty): row p's
    // columns [h_rf[p], h_rl[p]], column q's rows [h_cf[q], h_cl[q]]
    std::vector<int32_t> h_rf, h_rl, h_cf, h_cl;
    bool kd_ready = false;         // dK/domega0's band exists (adjoint)
    double* d_rkd = nullptr;       // dK/domega0 1
    double* d_ckd = nullptr;       // (dK/domega0)^T 1
    void* adj = nullptr;           // adjoint work (tg_adjoint.hip)
    bool assembled = false;

    // energy order: K is stored with rows and columns sorted by node energy
    int32_t* d_perm = nullptr;     // sorted position -> node
    std::vector<int32_t> h_perm;
    double* d_V_s = nullptr;       // V in sorted order
    double* d_eps_s = nullptr;     // eps in sorted order (energy fluxes)
    double* d_bath = nullptr;      // per-member bath occupations of the plain FP64 evaluations
                                   // (max_batch, device)
    // time-dependent bath (tg_set_bath_schedule): batch x ns samples
    std::vector<double> h_bsched;
    double* d_bsched = nullptr;
    int32_t bs_ns = 0, bs_batch = 0;
    double bs_t0 = 0, bs_dt = 0, bs_atol = 0, bs_rtol = 0;
    double* d_rk_s = nullptr;      // K 1 in sorted order
    double* d_ck_s = nullptr;      // K^T 1 in sorted order
    double* d_fs = nullptr;        // scratch: one state gathered to sorted order
    int64_t band_elems = 0;        // entries the plain evaluation streams (K's band)

    // FP64 band of K for the plain evaluation (tg_stack.hip)
    void* stack = nullptr;

    // mixed batched data (tg_mixed.hip)
    void* mixed = nullptr;

    // DP5 work (tg_dp5.hip)
    void* dp5 = nullptr;

    // electron-electron channel (tg_ee.hip), off when ee_g == 0
    double ee_g = 0, ee_kappa2 = 0, ee_sigma = 0, ee_dw = 0;
    void* ee = nullptr;

    // scratch vectors for host<->device staging (n x max_batch each)
    double* d_f = nullptr;
    double* d_I = nullptr;
    double* d_red = nullptr;       // reduction partials
    double* h_pinned = nullptr;    // pinned host scalars
    int32_t red_cap = 0;

    hipEvent_t ev_k0 = nullptr, ev_k1 = nullptr;  // dominant-kernel timing
    bool time_kernel = false;
    bool timed = false;

    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    int device = 0;
    std::string err;
};

namespace tg {

tg_status set_error(tg_ctx* c, tg_status code, const char* fmt, ...);

#define TG_HIP(ctx, call)                                                              \
    do {                                                                               \
        hipError_t _e = (call);                                                        \
        if (_e != hipSuccess)                                                          \
            return ::tg::set_error((ctx), TG_ERR_HIP, "%s failed: %s (%s:%d)", #call,  \
                                   hipGetErrorString(_e), __FILE__, __LINE__);         \
    } while (0)

#define TG_CHECK_LAUNCH(ctx) TG_HIP(ctx, hipGetLastError())

#define TG_TRY(expr)                          \
    do {                                      \
        tg_status _s = (expr);                \
        if (_s != TG_OK) return _s;           \
    } while (0)

template <typename T>
tg_status dev_alloc(tg_ctx* c, T** p, size_t count) {
    if (*p) return TG_OK;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T) + 16);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_error(c, TG_ERR_NOMEM, "hipMalloc(%zu bytes): %s", count * sizeof(T),
                         hipGetErrorString(e));
    }
    return TG_OK;
}

template <typename T>
void dev_free(T** p) {
    if (*p) hipFree((void*)*p);
    *p = nullptr;
}

inline unsigned ceil_div(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

// ---- internal entry points between translation units
// one plain FP64 evaluation; the bath is d_bath[0] (device) when given, else
// bath_n; d_flux (device, nullable) receives sum_i V_i eps_i I_i
tg_status plain_apply_device(tg_ctx* c, const double* d_f, double* d_I, double bath_n,
                             int conserve, const double* d_bath = nullptr, double* d_flux = nullptr);
// K's entries from the projection data with the convolved windows Q (the
// context's d_Q for K, the derivative table's for dK/domega0)
KSrc ksrc_of(const tg_ctx* c, const double* Q);
// exact nonzero extents of K in energy order (rows: first / last column;
// columns: first / last row; -1 when empty) and optionally the nonzero
// count; with split_scale > 0 of the fp16 hi/lo split at that scale
tg_status k_extents(tg_ctx* c, double split_scale, std::vector<int32_t>& rf, std::vector<int32_t>& rl,
                    std::vector<int32_t>& cf, std::vector<int32_t>& cl, int64_t* nnz);
tg_status assemble_domega(tg_ctx* c);
tg_status conservation_device(tg_ctx* c, const double* d_f, int32_t batch, double* number,
                              double* energy);
tg_status relative_error_device(tg_ctx* c, const double* d_err, const double* d_pop,
                                int32_t batch, double rtol, double atol, double* h_out);
// FP64 band of K (tg_stack.hip): built at assembly from the row extents
// h_rf / h_rl and the entries, one plain evaluation per call
tg_status stack_prepare(tg_ctx* c);
// dK/domega0 in the same band layout (entries from Qd)
tg_status stack_prepare_domega(tg_ctx* c, const double* d_Qd);
// u = M f, v = M^T f (caller order) from the same pass over M = K, or
// M = dK/domega0 when domega
tg_status stack_products_device(tg_ctx* c, const double* d_f, double* d_u, double* d_v, bool domega = false);
void stack_destroy(tg_ctx* c);
// records this device's address of the allocation-free evaluation's scratch
// array (tg_collision_apply_noalloc); called at context creation
tg_status noalloc_init(tg_ctx* c);
tg_status stack_apply_device(tg_ctx* c, const double* d_f, double* d_I, double bath_n, int conserve,
                             const double* d_bath, double* d_flux);
// from the block sums dpart (4 per block): with conserve I -= s w
// (number-defect removal; w = |f(1-f)|, or the hat form's weights d_wt when
// given); with flux the energy flux of the result
tg_status conserve_fix_device(tg_ctx* c, const double* dpart, int nblk, const double* d_f, const double* d_wt,
                              double* d_I, int conserve, double* flux);
tg_status mixed_prepare(tg_ctx* c);
tg_status mixed_apply_device(tg_ctx* c, const double* d_f, double* d_I, int32_t batch,
                             const double* h_bath, int conserve);
tg_status mixed_set_bath(tg_ctx* c, const double* h_bath, int32_t batch);
// gate (nullable, device, one per member): members with gate[b] == 0 are
// skipped (their outputs are not written)
// flux (nullable, device, one per member) receives each member's e-ph
// electron energy flux sum_i V_i eps_i I_i
tg_status mixed_eval_device(tg_ctx* c, const double* d_f, double* d_I, int32_t batch, int conserve,
                            const double* gate, double* flux = nullptr);
// Fused passes of the batched DP5 on the mixed path (available when member
// rows fit the LDS and there is no e-e channel): mixed_dp5_pass runs one
// Dp5PassArgs pass (tg_dp5_kernels.hpp) for every member, mixed_dp5_gemm the
// GEMM of the operands the last pass packed; kernel launches only
struct Dp5PassArgs;
bool mixed_dp5_fusable(tg_ctx* c);
tg_status mixed_dp5_pass(tg_ctx* c, int mode, const Dp5PassArgs& a, int32_t batch);
tg_status mixed_dp5_gemm(tg_ctx* c);
// before the passes of an attempt: running members (h[b] != 0) take the
// leading columns, and the GEMM computes only the member tiles they fill
tg_status mixed_dp5_compact(tg_ctx* c, const double* h, int32_t batch);
// the device array of per-member bath occupations the batched epilogue reads
// (a DP5 with a bath schedule writes it per stage; the host copy is then stale)
double* mixed_bath_device(tg_ctx* c);
void mixed_bath_forget(tg_ctx* c);
// error-norm sums sum_e (sum r^2 + (sum r)^2) per member into out[batch]
// (part: batch x relerr_blocks partials); kernel launches only
int relerr_blocks(tg_ctx* c);
tg_status relerr_sums_device(tg_ctx* c, const double* d_err, const double* d_pop, int32_t batch,
                             double rtol, double atol, double* part, double* out);
void mixed_destroy(tg_ctx* c);
void mixed_stats(tg_ctx* c, tg_kernel_stats* out);
void dp5_destroy(tg_ctx* c);
tg_status ee_prepare(tg_ctx* c, const double* h_eps);
void ee_destroy(tg_ctx* c);
// d_I[b] += I_ee(d_f[b]); no-op without the e-e channel; kernel launches only
tg_status ee_apply_device(tg_ctx* c, const double* d_f, double* d_I, int32_t batch,
                          const double* gate = nullptr);
// d_out += J_ee(f)^T mu for one member (nodes): the e-e operator's
// vector-Jacobian product, for the adjoint (tg_ee.hip)
tg_status ee_vjp_device(tg_ctx* c, const double* d_f, const double* d_mu, double* d_out);
void adjoint_destroy(tg_ctx* c);
tg_status ensure_staging(tg_ctx* c, int32_t batch);
// event bracket of the dominant kernel when TG_TIME_KERNEL is active
void time_begin(tg_ctx* c);
void time_end(tg_ctx* c);

}  // namespace tg
