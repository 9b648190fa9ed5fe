// tg_assemble.hip -- one-time FP64 assembly of the emission kernel K.
//
//   K_ij = g^2 sum_{b,b'} R_ib R_jb' G_sigma((b-b') dE + omega0)
//
// R_ib = int phi_i(k) h_b(eps(k)) dk projects node i's P1 basis function on
// the hat functions h_b of a uniform energy grid (eps linear per simplex,
// as the reference's LinearForm dispersion, Band.cpp:43-56).  Per simplex
// the basis-weighted energy density is vol * N(x)/span, N the order-(dim+1)
// B-spline on the sorted knots {e_0..e_dim, e_a} (simplex-spline identity),
// integrated exactly against the hats with 3-point Gauss-Legendre on every
// piece between knots and grid points.  The amplitude convention (g^2, full
// momentum measure, no 2*pi) follows ScatteringChannel.cpp:399-415.
//
// K itself is never stored as an n x n array: its entries
// K_ij = g^2 <R_i, Q_j> (tg_kentry.hpp) are evaluated by the kernels that
// build the bands the evaluations stream (tg_stack.hip for FP64,
// tg_mixed.hip for the fp16 split), from R and Q, which stay resident
// (n x maxw and n x (maxw + nm - 1) doubles).
//
// Kernels (all setup, run once per context):
//   proj_window_kernel  node -> [off, len) of touched hats        (gather)
//   proj_values_kernel  node -> R row                              (gather)
//   conv_kernel         node -> Q_j = R_j * g (sampled Gaussian)   (window conv)
//   k_extent_kernel     row p -> exact nonzero extent of K's energy-ordered
//                       row and column, nonzero count (entries where the
//                       windows meet)
//   K 1 and K^T 1 come from the FP64 band pass with f = 1
#include <algorithm>
#include <cmath>
#include <vector>

#include "tg_internal.hpp"
#include "tg_mesh.hpp"

namespace tg {

struct EGrid {
    double e0, de;
    int nbins;
};

__host__ __device__ inline int egrid_cell(const EGrid& g, double x) {
    int c = (int)floor((x - g.e0) / g.de);
    if (c < 0) c = 0;
    if (c > g.nbins - 2) c = g.nbins - 2;
    return c;
}

__device__ inline void sort_small(double* v, int n) {
    for (int i = 1; i < n; ++i) {
        double x = v[i];
        int j = i - 1;
        while (j >= 0 && v[j] > x) {
            v[j + 1] = v[j];
            --j;
        }
        v[j + 1] = x;
    }
}

// Cox-de Boor: order-k B-spline N_{0,k} on knots t[0..k], x inside span s.
__device__ inline double bspline_value(const double* t, int k, int s, double x) {
    double N[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    N[s] = 1.0;
    for (int q = 2; q <= k; ++q) {
        for (int i = 0; i <= k - q; ++i) {
            double v = 0.0;
            double d1 = t[i + q - 1] - t[i];
            double d2 = t[i + q] - t[i + 1];
            if (d1 > 0.0) v += (x - t[i]) / d1 * N[i];
            if (d2 > 0.0) v += (t[i + q] - x) / d2 * N[i + 1];
            N[i] = v;
        }
    }
    return N[0];
}

__device__ __constant__ double kGL3X[3] = {0.11270166537925831148, 0.5, 0.88729833462074168852};
__device__ __constant__ double kGL3W[3] = {0.27777777777777777778, 0.44444444444444444444,
                                           0.27777777777777777778};

// Adds the projection of simplex (energies ev) basis a onto the hats into
// acc[b - base].
__device__ void project_simplex_basis(int dim, const double* ev, int a, double vol,
                                      const EGrid& g, double* acc, int base) {
    double t[5];
    for (int q = 0; q <= dim; ++q) t[q] = ev[q];
    t[dim + 1] = ev[a];
    sort_small(t, dim + 2);
    const int nk = dim + 2;
    const double span = t[nk - 1] - t[0];
    const double de = g.de;
    if (!(span > 1e-9 * de)) {
        double x = t[0];
        int c = egrid_cell(g, x);
        double ec = g.e0 + (double)c * de;
        double w1 = (x - ec) / de;
        double m = vol / (double)(dim + 1);
        acc[c - base] += m * (1.0 - w1);
        acc[c + 1 - base] += m * w1;
        return;
    }
    const double scale = vol / span;
    const int c0 = egrid_cell(g, t[0]), c1 = egrid_cell(g, t[nk - 1]);
    for (int c = c0; c <= c1; ++c) {
        double ec = g.e0 + (double)c * de;
        double ec1 = ec + de;
        double lo = t[0] > ec ? t[0] : ec;
        double hi = t[nk - 1] < ec1 ? t[nk - 1] : ec1;
        if (!(hi > lo)) continue;
        for (int s = 0; s < nk - 1; ++s) {
            double a0 = t[s] > lo ? t[s] : lo;
            double a1 = t[s + 1] < hi ? t[s + 1] : hi;
            if (!(a1 > a0)) continue;
            double w = a1 - a0;
            double s0 = 0.0, s1 = 0.0;
            for (int q = 0; q < 3; ++q) {
                double x = a0 + kGL3X[q] * w;
                double v = kGL3W[q] * w * scale * bspline_value(t, nk - 1, s, x);
                s0 += v * (ec1 - x);
                s1 += v * (x - ec);
            }
            acc[c - base] += s0 / de;
            acc[c + 1 - base] += s1 / de;
        }
    }
}

__global__ void proj_window_kernel(MeshDims md, const double* __restrict__ eps,
                                   const int32_t* __restrict__ conn, EGrid g, int64_t n,
                                   int32_t* __restrict__ off, int32_t* __restrict__ len) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t elems[24];
    int locs[24];
    int ns = node_star(md, i, elems, locs);
    int lo = 0x7fffffff, hi = -1;
    for (int q = 0; q < ns; ++q) {
        double emin = 1e300, emax = -1e300;
        for (int a = 0; a <= md.dim; ++a) {
            double e = eps[conn[elems[q] * (md.dim + 1) + a]];
            emin = e < emin ? e : emin;
            emax = e > emax ? e : emax;
        }
        int c0 = egrid_cell(g, emin), c1 = egrid_cell(g, emax) + 1;
        lo = c0 < lo ? c0 : lo;
        hi = c1 > hi ? c1 : hi;
    }
    off[i] = lo;
    len[i] = hi - lo + 1;
}

__global__ void proj_values_kernel(MeshDims md, const double* __restrict__ eps,
                                   const int32_t* __restrict__ conn, EGrid g, double vol,
                                   int64_t n, const int32_t* __restrict__ off, int maxw,
                                   double* __restrict__ R) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double* row = R + i * (int64_t)maxw;
    for (int w = 0; w < maxw; ++w) row[w] = 0.0;
    int64_t elems[24];
    int locs[24];
    int ns = node_star(md, i, elems, locs);
    for (int q = 0; q < ns; ++q) {
        double ev[4];
        for (int a = 0; a <= md.dim; ++a) ev[a] = eps[conn[elems[q] * (md.dim + 1) + a]];
        project_simplex_basis(md.dim, ev, locs[q], vol, g, row, off[i]);
    }
}

// Q_j[p] = sum_w R_j[w] g[p - w]  (window of node j shifted by m_lo)
__global__ void conv_kernel(const double* __restrict__ R, const int32_t* __restrict__ len,
                            int maxw, const double* __restrict__ gm, int nm, int qw, int64_t n,
                            double* __restrict__ Q) {
    int64_t j = blockIdx.y;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n || p >= qw) return;
    const int lenj = len[j];
    const double* Rj = R + j * (int64_t)maxw;
    double s = 0.0;
    if (p < lenj + nm - 1) {
        int w0 = p - (nm - 1) > 0 ? p - (nm - 1) : 0;
        int w1 = p < lenj - 1 ? p : lenj - 1;
        for (int w = w0; w <= w1; ++w) s += Rj[w] * gm[p - w];
    }
    Q[j * (int64_t)qw + p] = s;
}

// Exact nonzero extents of K in energy order, from the entries themselves:
// block p takes row p (node perm[p], its window in LDS), tests every column's
// window and evaluates the entries where the windows meet; an entry counts
// when it is nonzero in FP64 or, with split_scale > 0, when its fp16 hi/lo
// split at that scale is (the batched path's band).  Row p gets its first
// and last such column and their count, column q its first and last row
// (atomic min / max: the result does not depend on their order).
constexpr int kExtT = 256;
__global__ void __launch_bounds__(kExtT)
k_extent_kernel(KSrc k, const int32_t* __restrict__ perm, int64_t n, double split_scale,
                int32_t* __restrict__ rf, int32_t* __restrict__ rl, int32_t* __restrict__ cf,
                int32_t* __restrict__ cl, unsigned long long* __restrict__ cnt) {
    extern __shared__ double sRi[];
    __shared__ int32_t sf[kExtT / 64], sl[kExtT / 64];
    __shared__ unsigned long long sc[kExtT / 64];
    const int64_t p = blockIdx.x;
    const int64_t i = perm[p];
    const int offi = k.off[i], leni = k.len[i];
    for (int w = threadIdx.x; w < leni; w += blockDim.x) sRi[w] = k.R[i * (int64_t)k.maxw + w];
    __syncthreads();
    int32_t f = INT32_MAX, l = -1;
    unsigned long long cc = 0;
    for (int64_t q = threadIdx.x; q < n; q += blockDim.x) {
        const int64_t j = perm[q];
        if (!k_windows_meet(k, offi, leni, j)) continue;
        const double v = k_entry_w(k, sRi, offi, leni, j);
        const bool nz = split_scale > 0.0 ? split_nonzero(v, split_scale) : v != 0.0;
        if (!nz) continue;
        f = min(f, (int32_t)q);
        l = (int32_t)q;
        ++cc;
        atomicMin(&cf[q], (int32_t)p);
        atomicMax(&cl[q], (int32_t)p);
    }
    for (int o = 32; o > 0; o >>= 1) {
        f = min(f, __shfl_down(f, o, 64));
        l = max(l, __shfl_down(l, o, 64));
        cc += __shfl_down(cc, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sf[threadIdx.x >> 6] = f;
        sl[threadIdx.x >> 6] = l;
        sc[threadIdx.x >> 6] = cc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kExtT / 64; ++w) {
            f = min(f, sf[w]);
            l = max(l, sl[w]);
            cc += sc[w];
        }
        rf[p] = l < 0 ? -1 : f;
        rl[p] = l;
        if (cnt) cnt[p] = cc;
    }
}

__global__ void fill_i32_kernel(int32_t* __restrict__ a, int64_t n, int32_t v) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void fill_f64_kernel(double* __restrict__ a, int64_t n, double v) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// natural-order rows of K from the entries: out[r][j] = K_{row0 + r, j}
__global__ void copy_rows_kernel(KSrc k, int64_t n, int64_t row0, double* __restrict__ out) {
    extern __shared__ double sRi[];
    const int64_t r = blockIdx.x;
    const int64_t i = row0 + r;
    const int offi = k.off[i], leni = k.len[i];
    for (int w = threadIdx.x; w < leni; w += blockDim.x) sRi[w] = k.R[i * (int64_t)k.maxw + w];
    __syncthreads();
    for (int64_t j = threadIdx.x; j < n; j += blockDim.x)
        out[r * n + j] = k_windows_meet(k, offi, leni, j) ? k_entry_w(k, sRi, offi, leni, j) : 0.0;
}

__global__ void gather_kernel(const double* __restrict__ src, const int32_t* __restrict__ perm,
                              int64_t n, double* __restrict__ dst) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p < n) dst[p] = src[perm[p]];
}

// K's entries with the convolved windows Q
KSrc ksrc_of(const tg_ctx* c, const double* Q) {
    return KSrc{c->d_R, Q, c->d_off, c->d_len, c->maxw, c->qw, c->m_lo, c->nm, c->g * c->g};
}

// Exact extents of K in energy order (host copies), optionally of the fp16
// split at split_scale, and the nonzero count; kernel scan over all pairs
// whose windows meet.
tg_status k_extents(tg_ctx* c, double split_scale, std::vector<int32_t>& rf, std::vector<int32_t>& rl,
                    std::vector<int32_t>& cf, std::vector<int32_t>& cl, int64_t* nnz) {
    const int64_t n = c->n;
    int32_t *d_rf = nullptr, *d_rl = nullptr, *d_cf = nullptr, *d_cl = nullptr;
    unsigned long long* d_cnt = nullptr;
    TG_TRY(dev_alloc(c, &d_rf, (size_t)n));
    TG_TRY(dev_alloc(c, &d_rl, (size_t)n));
    TG_TRY(dev_alloc(c, &d_cf, (size_t)n));
    TG_TRY(dev_alloc(c, &d_cl, (size_t)n));
    if (nnz) TG_TRY(dev_alloc(c, &d_cnt, (size_t)n));
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(fill_i32_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, d_cf, n, INT32_MAX);
    hipLaunchKernelGGL(fill_i32_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, d_cl, n, -1);
    hipLaunchKernelGGL(k_extent_kernel, dim3((unsigned)n), dim3(kExtT), sizeof(double) * c->maxw, st,
                       ksrc_of(c, c->d_Q), c->d_perm, n, split_scale, d_rf, d_rl, d_cf, d_cl, d_cnt);
    TG_CHECK_LAUNCH(c);
    rf.resize(n);
    rl.resize(n);
    cf.resize(n);
    cl.resize(n);
    TG_HIP(c, hipMemcpyAsync(rf.data(), d_rf, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    TG_HIP(c, hipMemcpyAsync(rl.data(), d_rl, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    TG_HIP(c, hipMemcpyAsync(cf.data(), d_cf, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    TG_HIP(c, hipMemcpyAsync(cl.data(), d_cl, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    std::vector<unsigned long long> cnt(nnz ? n : 0);
    if (nnz) TG_HIP(c, hipMemcpyAsync(cnt.data(), d_cnt, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost, st));
    TG_HIP(c, hipStreamSynchronize(st));
    for (int64_t q = 0; q < n; ++q)
        if (cl[q] < 0) cf[q] = -1;
    if (nnz) {
        *nnz = 0;
        for (auto v : cnt) *nnz += (int64_t)v;
    }
    dev_free(&d_rf);
    dev_free(&d_rl);
    dev_free(&d_cf);
    dev_free(&d_cl);
    dev_free(&d_cnt);
    return TG_OK;
}

}  // namespace tg

using namespace tg;

extern "C" tg_status tg_kernel_assemble(tg_ctx* c, int32_t flags) {
    if (!c) return TG_ERR_ARG;
    MeshDims md{c->dim, {c->nsec[0], c->nsec[1], c->nsec[2]}};
    EGrid g{c->e0, c->de, c->nbins};
    const unsigned nb = ceil_div(c->n, 128);

    if (!c->assembled) {
        // 1. projection windows
        TG_TRY(dev_alloc(c, &c->d_off, c->n));
        TG_TRY(dev_alloc(c, &c->d_len, c->n));
        hipLaunchKernelGGL(proj_window_kernel, dim3(nb), dim3(128), 0, c->stream, md, c->d_eps,
                           c->d_conn, g, c->n, c->d_off, c->d_len);
        TG_CHECK_LAUNCH(c);
        std::vector<int32_t> hlen(c->n);
        TG_HIP(c, hipMemcpyAsync(hlen.data(), c->d_len, sizeof(int32_t) * c->n,
                                 hipMemcpyDeviceToHost, c->stream));
        TG_HIP(c, hipStreamSynchronize(c->stream));
        int32_t maxw = 0;
        for (auto v : hlen) maxw = v > maxw ? v : maxw;
        c->maxw = maxw;

        // 2. projection values
        TG_TRY(dev_alloc(c, &c->d_R, (size_t)c->n * maxw));
        hipLaunchKernelGGL(proj_values_kernel, dim3(nb), dim3(128), 0, c->stream, md, c->d_eps,
                           c->d_conn, g, c->vol, c->n, c->d_off, maxw, c->d_R);
        TG_CHECK_LAUNCH(c);

        // 3. sampled broadened delta G(m dE + omega0), |.| <= cutoff*sigma
        const int32_t lo = (int32_t)std::ceil((-c->cutoff * c->sigma - c->omega0) / c->de);
        const int32_t hi = (int32_t)std::floor((c->cutoff * c->sigma - c->omega0) / c->de);
        c->m_lo = lo;
        c->nm = hi - lo + 1;
        std::vector<double> gm(c->nm);
        const double norm = 1.0 / (c->sigma * std::sqrt(2.0 * M_PI));
        for (int32_t m = lo; m <= hi; ++m) {
            double x = (double)m * c->de + c->omega0;
            gm[m - lo] = norm * std::exp(-0.5 * (x / c->sigma) * (x / c->sigma));
        }
        TG_TRY(dev_alloc(c, &c->d_gm, gm.size()));
        TG_HIP(c, hipMemcpyAsync(c->d_gm, gm.data(), sizeof(double) * gm.size(),
                                 hipMemcpyHostToDevice, c->stream));

        // 4. Q = R (*) g
        c->qw = c->maxw + c->nm - 1;
        TG_TRY(dev_alloc(c, &c->d_Q, (size_t)c->n * c->qw));
        hipLaunchKernelGGL(conv_kernel, dim3(ceil_div(c->qw, 128), c->n), dim3(128), 0, c->stream,
                           c->d_R, c->d_len, c->maxw, c->d_gm, c->nm, c->qw, c->n, c->d_Q);
        TG_CHECK_LAUNCH(c);

        // 5. exact nonzero extents of K in energy order and its nonzero count
        TG_TRY(k_extents(c, 0.0, c->h_rf, c->h_rl, c->h_cf, c->h_cl, &c->nnz));
        // 6. the FP64 band of K of the plain evaluation, entries from R and Q
        TG_TRY(stack_prepare(c));
        // 7. K 1 and K^T 1 through the same band pass (f = 1)
        TG_TRY(dev_alloc(c, &c->d_rk, c->n));
        TG_TRY(dev_alloc(c, &c->d_ck, c->n));
        TG_TRY(dev_alloc(c, &c->d_rk_s, c->n));
        TG_TRY(dev_alloc(c, &c->d_ck_s, c->n));
        {
            double* d_one = nullptr;
            TG_TRY(dev_alloc(c, &d_one, (size_t)c->n));
            hipLaunchKernelGGL(fill_f64_kernel, dim3(ceil_div(c->n, 256)), dim3(256), 0, c->stream, d_one, c->n,
                               1.0);
            tg_status sp = stack_products_device(c, d_one, c->d_rk, c->d_ck);
            if (sp == TG_OK) {
                hipLaunchKernelGGL(gather_kernel, dim3(ceil_div(c->n, 256)), dim3(256), 0, c->stream,
                                   c->d_rk, c->d_perm, c->n, c->d_rk_s);
                hipLaunchKernelGGL(gather_kernel, dim3(ceil_div(c->n, 256)), dim3(256), 0, c->stream,
                                   c->d_ck, c->d_perm, c->n, c->d_ck_s);
                if (hipGetLastError() != hipSuccess) sp = set_error(c, TG_ERR_HIP, "row sums of K");
            }
            if (sp == TG_OK && hipStreamSynchronize(c->stream) != hipSuccess)
                sp = set_error(c, TG_ERR_HIP, "row sums of K");
            dev_free(&d_one);
            TG_TRY(sp);
        }
        c->assembled = true;
    }
    if ((flags & TG_PREC_MIXED) || c->max_batch > 1) TG_TRY(mixed_prepare(c));
    return TG_OK;
}

extern "C" tg_status tg_kernel_copy_rows(tg_ctx* c, int64_t row0, int64_t nrows, double* out) {
    if (!c || !out) return TG_ERR_ARG;
    if (!c->assembled) return set_error(c, TG_ERR_STATE, "kernel not assembled");
    if (row0 < 0 || nrows < 0 || row0 + nrows > c->n) return set_error(c, TG_ERR_ARG, "rows out of range");
    if (nrows == 0) return TG_OK;
    // rows computed from the projection data (K is never stored densely)
    double* d_out = nullptr;
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(nrows, (int64_t)(1ll << 27) / c->n));
    TG_TRY(dev_alloc(c, &d_out, (size_t)chunk * c->n));
    const KSrc k = ksrc_of(c, c->d_Q);
    hipError_t e = hipSuccess;
    for (int64_t r0 = 0; r0 < nrows && e == hipSuccess; r0 += chunk) {
        const int64_t nr = std::min(chunk, nrows - r0);
        hipLaunchKernelGGL(copy_rows_kernel, dim3((unsigned)nr), dim3(256), sizeof(double) * c->maxw, c->stream, k,
                           c->n, row0 + r0, d_out);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(out + r0 * c->n, d_out, sizeof(double) * nr * c->n,
                               hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    dev_free(&d_out);
    if (e != hipSuccess) return set_error(c, TG_ERR_HIP, "copy rows: %s", hipGetErrorString(e));
    return TG_OK;
}

namespace tg {

// dK/domega0: the same projection windows with the sampled derivative table
// G'(x) = -x/sigma^2 G(x) (x = m dE + omega0), then K' 1 and K'^T 1.
tg_status assemble_domega(tg_ctx* c) {
    if (c->kd_ready) return TG_OK;
    if (!c->assembled) return set_error(c, TG_ERR_STATE, "kernel not assembled");
    std::vector<double> gd(c->nm);
    const double norm = 1.0 / (c->sigma * std::sqrt(2.0 * M_PI));
    for (int32_t m = c->m_lo; m < c->m_lo + c->nm; ++m) {
        const double x = (double)m * c->de + c->omega0;
        gd[m - c->m_lo] = -x / (c->sigma * c->sigma) * norm *
                          std::exp(-0.5 * (x / c->sigma) * (x / c->sigma));
    }
    double* d_gd = nullptr;
    double* d_Qd = nullptr;
    double* d_one = nullptr;
    TG_TRY(dev_alloc(c, &d_gd, gd.size()));
    TG_TRY(dev_alloc(c, &d_Qd, (size_t)c->n * c->qw));
    TG_TRY(dev_alloc(c, &d_one, (size_t)c->n));
    TG_HIP(c, hipMemcpyAsync(d_gd, gd.data(), sizeof(double) * gd.size(), hipMemcpyHostToDevice,
                             c->stream));
    hipLaunchKernelGGL(conv_kernel, dim3(ceil_div(c->qw, 128), c->n), dim3(128), 0, c->stream,
                       c->d_R, c->d_len, c->maxw, d_gd, c->nm, c->qw, c->n, d_Qd);
    TG_CHECK_LAUNCH(c);
    // dK/domega0 in K's band layout (its nonzeros lie inside K's: the same
    // windows, and a window pair whose K entry is 0 has Q_j = 0 on the overlap)
    TG_TRY(stack_prepare_domega(c, d_Qd));
    TG_TRY(dev_alloc(c, &c->d_rkd, c->n));
    TG_TRY(dev_alloc(c, &c->d_ckd, c->n));
    hipLaunchKernelGGL(fill_f64_kernel, dim3(ceil_div(c->n, 256)), dim3(256), 0, c->stream, d_one, c->n, 1.0);
    TG_CHECK_LAUNCH(c);
    TG_TRY(stack_products_device(c, d_one, c->d_rkd, c->d_ckd, true));
    TG_HIP(c, hipStreamSynchronize(c->stream));
    dev_free(&d_one);
    dev_free(&d_Qd);
    dev_free(&d_gd);
    c->kd_ready = true;
    return TG_OK;
}

}  // namespace tg

extern "C" tg_status tg_kernel_stats_get(tg_ctx* c, tg_kernel_stats* out) {
    if (!c || !out) return TG_ERR_ARG;
    if (!c->assembled) return set_error(c, TG_ERR_STATE, "kernel not assembled");
    *out = tg_kernel_stats{};
    out->n = c->n;
    out->nnz = c->nnz;
    out->plain_band = c->band_elems;
    out->window = c->maxw;
    std::vector<double> rk(c->n);
    TG_HIP(c, hipMemcpyAsync(rk.data(), c->d_rk, sizeof(double) * c->n, hipMemcpyDeviceToHost,
                             c->stream));
    TG_HIP(c, hipStreamSynchronize(c->stream));
    for (double v : rk) out->kmax = v > out->kmax ? v : out->kmax;
    mixed_stats(c, out);
    return TG_OK;
}
