// tg_plain.hip -- FP64 products u = M f and v = M^T f from one pass over a
// band matrix M stored n x ld (row-major, energy order): the row and column
// sums of K at assembly and the products with dK/domega0 in the adjoint.  The
// collision operator's own FP64 evaluation streams K's compact band
// (tg_stack.hip); this file also holds the C entry point tg_collision_apply
// and the number-defect / energy-flux reduction its finish kernels share.
//
// K is stored with rows and columns in energy order, which makes the
// emission kernel a band: every 128-row block has a contiguous range of
// nonzero columns, and only (row block, 2048-column chunk) pairs that meet
// it are launched; sweeps of a launched chunk outside the range publish
// zeros without reading M.
//
//   phase 1 (plain_dual_kernel): workgroup (row block r, chunk c) covers
//     rows [128 r, 128 r + 128) x cols [2048 c, 2048 c + 2048).  Each of the 4
//     waves owns 32 consecutive rows; a lane owns two adjacent columns of a
//     128-column sweep, so every row read is one 1 KiB coalesced segment per
//     wave (16 B per lane).  The 32 loads of a sweep are independent and in
//     flight together.  Row partials of u stay in registers for the whole
//     chunk and are reduced across the wave once per chunk; column partials
//     of v are reduced across the 4 waves in LDS once per sweep.
//       upart[c][i] = sum_{j in chunk c} M_ij f_j
//       vpart[r][j] = sum_{i in rowblock r} M_ij f_i
//   phase 2 (plain_finish_kernel): per node, sum the partials in index order
//     (deterministic) and write u and v in the caller's node order.
#include "tg_internal.hpp"

namespace tg {

typedef double dbl2 __attribute__((ext_vector_type(2)));

constexpr int kRowsPerWave = kPlainRows / 4;            // 32
constexpr int kSweeps = kPlainChunk / kPlainCols;       // 16

__global__ void __launch_bounds__(kPlainThreads)
plain_dual_kernel(const double* __restrict__ K, int64_t n, int64_t ld,
                  const double* __restrict__ f, const int2* __restrict__ work,
                  const int2* __restrict__ rb_col, double* __restrict__ upart,
                  double* __restrict__ vpart) {
    __shared__ double sv[4][kPlainCols];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int2 wk = work[blockIdx.x];           // (row block, column chunk)
    const int2 band = rb_col[wk.x];              // nonzero columns of the row block
    const int64_t row0 = (int64_t)wk.x * kPlainRows + wave * kRowsPerWave;
    const int64_t col0 = (int64_t)wk.y * kPlainChunk;

    double fr[kRowsPerWave];
    double uacc[kRowsPerWave];
#pragma unroll
    for (int q = 0; q < kRowsPerWave; ++q) {
        const int64_t r = row0 + q;
        fr[q] = r < n ? f[r] : 0.0;
        uacc[q] = 0.0;
    }
    const int nrows = (int)((n - row0) < kRowsPerWave ? (n - row0) : kRowsPerWave);

    for (int s = 0; s < kSweeps; ++s) {
        const int64_t sw = col0 + (int64_t)s * kPlainCols;
        if (sw >= ld) break;  // uniform over the block
        const bool in_band = sw >= band.x && sw < band.y;  // uniform over the block
        const int64_t cs = sw + 2 * lane;
        double v0 = 0.0, v1 = 0.0;
        // ld is a multiple of 64, not of 128: lanes past the row end stay idle
        if (in_band && nrows > 0 && cs < ld) {
            double fc0 = 0.0, fc1 = 0.0;
            if (cs < n) fc0 = f[cs];
            if (cs + 1 < n) fc1 = f[cs + 1];
            if (nrows == kRowsPerWave) {
                dbl2 kk[kRowsPerWave];
#pragma unroll
                for (int q = 0; q < kRowsPerWave; ++q)
                    kk[q] = __builtin_nontemporal_load(
                        reinterpret_cast<const dbl2*>(K + (row0 + q) * ld + cs));
#pragma unroll
                for (int q = 0; q < kRowsPerWave; ++q) {
                    uacc[q] += kk[q].x * fc0 + kk[q].y * fc1;
                    v0 += kk[q].x * fr[q];
                    v1 += kk[q].y * fr[q];
                }
            } else {
                for (int q = 0; q < nrows; ++q) {
                    dbl2 kk = *reinterpret_cast<const dbl2*>(K + (row0 + q) * ld + cs);
                    uacc[q] += kk.x * fc0 + kk.y * fc1;
                    v0 += kk.x * fr[q];
                    v1 += kk.y * fr[q];
                }
            }
        }
        // sweeps outside the band publish zeros (the chunk is covered by this
        // row block, so the finishing sum reads them)
        sv[wave][2 * lane] = v0;
        sv[wave][2 * lane + 1] = v1;
        __syncthreads();
        if (threadIdx.x < kPlainCols) {
            const int64_t cc = sw + threadIdx.x;
            const double t = (sv[0][threadIdx.x] + sv[1][threadIdx.x]) +
                             (sv[2][threadIdx.x] + sv[3][threadIdx.x]);
            if (cc < ld) vpart[(int64_t)wk.x * ld + cc] = t;
        }
        __syncthreads();
    }
    // row partials: reduce across the 64 lanes of the wave
#pragma unroll
    for (int q = 0; q < kRowsPerWave; ++q) {
        double u = uacc[q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) u += __shfl_xor(u, o, 64);
        uacc[q] = u;
    }
    if (lane < kRowsPerWave) {
        double mine = 0.0;
#pragma unroll
        for (int q = 0; q < kRowsPerWave; ++q) mine = (lane == q) ? uacc[q] : mine;
        const int64_t r = row0 + lane;
        if (r < n) upart[(int64_t)wk.y * n + r] = mine;
    }
}

__global__ void gather_sorted_f_kernel(const double* __restrict__ f, const int32_t* __restrict__ perm,
                                       int64_t n, double* __restrict__ fs) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p < n) fs[p] = f[perm[p]];
}

// Raw products u -> out0, v -> out1 (caller order perm[p]).  Thread (tx, ty)
// sums slice ty of node p's partials: the column chunks of its row block (u)
// and the row blocks covering its column chunk (v).
constexpr int kFinishNodes = 64;
constexpr int kFinishSlices = 4;

__global__ void __launch_bounds__(kFinishNodes* kFinishSlices)
plain_finish_kernel(const double* __restrict__ upart, const double* __restrict__ vpart,
                    const int2* __restrict__ rb_chunk, const int32_t* __restrict__ cc_off,
                    const int32_t* __restrict__ cc_rows, int64_t n, int64_t ld,
                    const int32_t* __restrict__ perm, double* __restrict__ out0,
                    double* __restrict__ out1) {
    __shared__ double su[kFinishSlices][kFinishNodes];
    __shared__ double svv[kFinishSlices][kFinishNodes];
    const int tx = threadIdx.x % kFinishNodes;
    const int ty = threadIdx.x / kFinishNodes;
    const int64_t p = (int64_t)blockIdx.x * kFinishNodes + tx;
    double u = 0.0, v = 0.0;
    if (p < n) {
        const int2 ch = rb_chunk[p / kPlainRows];
        for (int c = ch.x + ty; c <= ch.y; c += kFinishSlices) u += upart[(int64_t)c * n + p];
        const int cchunk = (int)(p / kPlainChunk);
        for (int q = cc_off[cchunk] + ty; q < cc_off[cchunk + 1]; q += kFinishSlices)
            v += vpart[(int64_t)cc_rows[q] * ld + p];
    }
    su[ty][tx] = u;
    svv[ty][tx] = v;
    __syncthreads();
    if (ty == 0 && p < n) {
        u = (su[0][tx] + su[1][tx]) + (su[2][tx] + su[3][tx]);
        v = (svv[0][tx] + svv[1][tx]) + (svv[2][tx] + svv[3][tx]);
        const int64_t node = perm[p];
        out0[node] = u;
        out1[node] = v;
    }
}

// Per-block partials dpart[4 b + {0,1,2,3}] = sums of V I, V |f(1-f)|,
// V eps I and V eps |f(1-f)| (fixed order) give
//   - with conserve, the number-defect removal I_i -= s |f_i(1-f_i)|,
//     s = sum V I / sum V |f(1-f)| (weights nonnegative also where a state
//     leaves [0,1]);
//   - with flux, the e-ph electron energy flux of the corrected I:
//     sum V eps I - s sum V eps |f(1-f)| (one value, written by thread 0).
__global__ void conserve_fix_kernel(const double* __restrict__ dpart, int nblk, int64_t n,
                                    const double* __restrict__ f, double* __restrict__ I,
                                    int conserve, double* __restrict__ flux) {
    double d = 0.0, w = 0.0, ei = 0.0, ew = 0.0;
    for (int b = 0; b < nblk; ++b) {
        d += dpart[4 * b];
        w += dpart[4 * b + 1];
        ei += dpart[4 * b + 2];
        ew += dpart[4 * b + 3];
    }
    const double s = (conserve && w > 0.0) ? d / w : 0.0;
    if (flux && blockIdx.x == 0 && threadIdx.x == 0) *flux = ei - s * ew;
    if (s == 0.0) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        I[i] -= s * fabs(f[i] * (1.0 - f[i]));
}

tg_status conserve_fix_device(tg_ctx* c, const double* dpart, int nblk, const double* d_f, double* d_I,
                              int conserve, double* flux) {
    hipLaunchKernelGGL(conserve_fix_kernel, dim3(conserve ? ceil_div(c->n, 256) : 1), dim3(256), 0, c->stream,
                       dpart, nblk, c->n, d_f, d_I, conserve, flux);
    TG_CHECK_LAUNCH(c);
    return TG_OK;
}

static tg_status launch_dual(tg_ctx* c, const double* Kmat, const double* d_f) {
    hipLaunchKernelGGL(gather_sorted_f_kernel, dim3(ceil_div(c->n, 256)), dim3(256), 0, c->stream,
                       d_f, c->d_perm, c->n, c->d_fs);
    TG_CHECK_LAUNCH(c);
    time_begin(c);
    hipLaunchKernelGGL(plain_dual_kernel, dim3(c->n_work), dim3(kPlainThreads), 0, c->stream, Kmat,
                       c->n, c->ld, c->d_fs, c->d_work, c->d_rb_col, c->d_upart, c->d_vpart);
    time_end(c);
    TG_CHECK_LAUNCH(c);
    return TG_OK;
}

// one plain FP64 evaluation: one pass over K's band (tg_stack.hip)
tg_status plain_apply_device(tg_ctx* c, const double* d_f, double* d_I, double bath_n,
                             int conserve, const double* d_bath, double* d_flux) {
    if (!c->assembled) return set_error(c, TG_ERR_STATE, "kernel not assembled");
    return stack_apply_device(c, d_f, d_I, bath_n, conserve, d_bath, d_flux);
}

// u = M f, v = M^T f (caller's node order) for M = K or dK/domega0 stored in
// energy order (row/column sums at assembly, products of the adjoint)
tg_status plain_products_device(tg_ctx* c, const double* Kmat, const double* d_f, double* d_u,
                                double* d_v) {
    TG_TRY(launch_dual(c, Kmat, d_f));
    const unsigned nb = ceil_div(c->n, kFinishNodes);
    hipLaunchKernelGGL(plain_finish_kernel, dim3(nb), dim3(kFinishNodes * kFinishSlices), 0,
                       c->stream, c->d_upart, c->d_vpart, c->d_rb_chunk, c->d_cc_off,
                       c->d_cc_rows, c->n, c->ld, c->d_perm, d_u, d_v);
    TG_CHECK_LAUNCH(c);
    return TG_OK;
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
