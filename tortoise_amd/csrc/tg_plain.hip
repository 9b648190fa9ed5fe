// tg_plain.hip -- the C entry point tg_collision_apply (plain FP64 or
// batched mixed-precision evaluations, host or device pointers) and the
// number-defect / energy-flux reduction the FP64 finish kernels share.  The
// FP64 evaluation itself streams K's compact band (tg_stack.hip).
#include "tg_internal.hpp"

namespace tg {

// Per-block partials dpart[4 b + {0,1,2,3}] = sums of V I, V w, V eps I
// and V eps w (fixed order), with weights w = |f(1-f)| (nodal factors) or
// the hat form's weights wt (caller order), give
//   - with conserve, the number-defect removal I_i -= s w_i,
//     s = sum V I / sum V w (weights nonnegative also where a state leaves
//     [0,1]);
//   - with flux, the e-ph electron energy flux of the corrected I:
//     sum V eps I - s sum V eps w (one value, written by thread 0).
constexpr int kFixT = 256;

__global__ void __launch_bounds__(kFixT)
conserve_fix_kernel(const double* __restrict__ dpart, int nblk, int64_t n,
                    const double* __restrict__ f, const double* __restrict__ wt,
                    double* __restrict__ I, int conserve, double* __restrict__ flux) {
    // every block sums the partials the same way: strided per thread, then
    // a fixed tree over the lanes and waves
    __shared__ double red[4][kFixT / 64];
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblk; b += kFixT)
#pragma unroll
        for (int q = 0; q < 4; ++q) a[q] += dpart[4 * b + q];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        for (int o = 32; o > 0; o >>= 1) a[q] += __shfl_down(a[q], o, 64);
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = a[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        a[q] = 0.0;
        for (int w = 0; w < kFixT / 64; ++w) a[q] += red[q][w];
    }
    const double d = a[0], w = a[1], ei = a[2], ew = a[3];
    const double s = (conserve && w > 0.0) ? d / w : 0.0;
    if (flux && blockIdx.x == 0 && threadIdx.x == 0) *flux = ei - s * ew;
    if (s == 0.0) return;
    for (int64_t i = blockIdx.x * (int64_t)kFixT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kFixT)
        I[i] -= s * (wt ? wt[i] : fabs(f[i] * (1.0 - f[i])));
}

tg_status conserve_fix_device(tg_ctx* c, const double* dpart, int nblk, const double* d_f, const double* d_wt,
                              double* d_I, int conserve, double* flux) {
    hipLaunchKernelGGL(conserve_fix_kernel, dim3(conserve ? ceil_div(c->n, kFixT) : 1), dim3(kFixT), 0, c->stream,
                       dpart, nblk, c->n, d_f, d_wt, d_I, conserve, flux);
    TG_CHECK_LAUNCH(c);
    return TG_OK;
}

// one plain FP64 evaluation: one pass over K's band (tg_stack.hip)
tg_status plain_apply_device(tg_ctx* c, const double* d_f, double* d_I, double bath_n,
                             int conserve, const double* d_bath, double* d_flux) {
    if (!c->assembled) return set_error(c, TG_ERR_STATE, "kernel not assembled");
    return stack_apply_device(c, d_f, d_I, bath_n, conserve, d_bath, d_flux);
}

}  // namespace tg

using namespace tg;

extern "C" tg_status tg_collision_apply(tg_ctx* c, const double* f, double* I, int64_t n,
                                        int32_t batch, const double* bath_n, int32_t flags) {
    if (!c || !f || !I) return TG_ERR_ARG;
    if (n != c->n)
        return set_error(c, TG_ERR_ARG, "n=%lld but the mesh has %lld nodes", (long long)n,
                         (long long)c->n);
    if (batch < 1) return set_error(c, TG_ERR_ARG, "batch < 1");
    if (!c->assembled) return set_error(c, TG_ERR_STATE, "tg_kernel_assemble not called");
    const bool dev = flags & TG_PTR_DEVICE;
    const int conserve = (flags & TG_CONSERVE) ? 1 : 0;
    c->time_kernel = (flags & TG_TIME_KERNEL) != 0;
    const double* d_f = f;
    double* d_I = I;
    if (!dev) {
        TG_TRY(ensure_staging(c, batch));
        TG_HIP(c, hipMemcpyAsync(c->d_f, f, sizeof(double) * n * batch, hipMemcpyHostToDevice,
                                 c->stream));
        d_f = c->d_f;
        d_I = c->d_I;
    }
    if (flags & TG_PREC_MIXED) {
        TG_TRY(mixed_apply_device(c, d_f, d_I, batch, bath_n, conserve));
    } else {
        for (int b = 0; b < batch; ++b) {
            const double nb = bath_n ? bath_n[b] : c->bath_n;
            TG_TRY(plain_apply_device(c, d_f + (int64_t)b * n, d_I + (int64_t)b * n, nb, conserve, nullptr,
                                      nullptr));
        }
    }
    c->time_kernel = false;
    TG_TRY(ee_apply_device(c, d_f, d_I, batch));
    if (!dev) {
        TG_HIP(c, hipMemcpyAsync(I, d_I, sizeof(double) * n * batch, hipMemcpyDeviceToHost,
                                 c->stream));
        TG_HIP(c, hipStreamSynchronize(c->stream));
    }
    return TG_OK;
}
