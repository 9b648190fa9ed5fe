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
// Kernels (all setup, run once per context):
//   proj_window_kernel  node -> [off, len) of touched hats        (gather)
//   proj_values_kernel  node -> R row                              (gather)
//   conv_kernel         node -> Q_j = R_j * g (sampled Gaussian)   (window conv)
//   kernel_entries      (i, j) -> K_ij = g^2 <R_i, Q_j>            (dense FP64)
//   K 1 and K^T 1 come from the plain one-pass kernel with f = 1
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

// K_ij = amp * sum_b R_i(b) Q_j(b) over the overlap of the two windows,
// stored in energy order: row p holds node i = perm[p], column q node
// j = perm[q].  Block = 256 columns of one row; the row's window is in LDS.
constexpr int kKThreads = 256;
__global__ void kernel_entries(const double* __restrict__ R, const double* __restrict__ Q,
                               const int32_t* __restrict__ off, const int32_t* __restrict__ len,
                               int maxw, int qw, int m_lo, int nm, int64_t n, int64_t ld,
                               const int32_t* __restrict__ perm, double amp,
                               double* __restrict__ K) {
    extern __shared__ double sRi[];
    const int64_t prow = blockIdx.y;
    const int64_t i = perm[prow];
    const int offi = off[i], leni = len[i];
    for (int w = threadIdx.x; w < leni; w += blockDim.x) sRi[w] = R[i * (int64_t)maxw + w];
    __syncthreads();
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= ld) return;
    double s = 0.0;
    if (q < n) {
        const int64_t j = perm[q];
        const int qoff = off[j] + m_lo;
        const int qlen = len[j] + nm - 1;
        int b0 = offi > qoff ? offi : qoff;
        int b1 = (offi + leni < qoff + qlen) ? offi + leni : qoff + qlen;
        const double* Qj = Q + j * (int64_t)qw;
        for (int b = b0; b < b1; ++b) s += sRi[b - offi] * Qj[b - qoff];
        s *= amp;
    }
    K[prow * ld + q] = s;  // padding columns hold exact zeros
}

// first / last nonzero column of every (sorted) row of K; -1 when empty
__global__ void row_nz_extent_kernel(const double* __restrict__ K, int64_t n, int64_t ld,
                                     int32_t* __restrict__ first, int32_t* __restrict__ last) {
    __shared__ int32_t sf[4], sl[4];
    const int64_t p = blockIdx.x;
    int32_t f = 0x7fffffff, l = -1;
    for (int64_t q = threadIdx.x; q < n; q += blockDim.x)
        if (K[p * ld + q] != 0.0) {
            f = q < f ? (int32_t)q : f;
            l = (int32_t)q;
        }
    for (int o = 32; o > 0; o >>= 1) {
        f = min(f, __shfl_down(f, o, 64));
        l = max(l, __shfl_down(l, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        sf[threadIdx.x >> 6] = f;
        sl[threadIdx.x >> 6] = l;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        first[p] = min(min(sf[0], sf[1]), min(sf[2], sf[3]));
        last[p] = max(max(sl[0], sl[1]), max(sl[2], sl[3]));
    }
}

__global__ void gather_kernel(const double* __restrict__ src, const int32_t* __restrict__ perm,
                              int64_t n, double* __restrict__ dst) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p < n) dst[p] = src[perm[p]];
}

// natural-order rows of K from the sorted storage: out[r][j] = K[pinv r'][pinv j]
__global__ void copy_rows_kernel(const double* __restrict__ K, int64_t n, int64_t ld,
                                 const int32_t* __restrict__ pinv, int64_t row0,
                                 double* __restrict__ out) {
    const int64_t r = blockIdx.y;
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= n) return;
    out[r * n + j] = K[(int64_t)pinv[row0 + r] * ld + pinv[j]];
}

// nonzero count per row of K (for throughput accounting)
__global__ void nnz_rows_kernel(const double* __restrict__ K, int64_t n, int64_t ld,
                                unsigned long long* __restrict__ out) {
    __shared__ unsigned long long sh[4];
    const int64_t i = blockIdx.x;
    unsigned long long c = 0;
    for (int64_t j = threadIdx.x; j < n; j += blockDim.x) c += K[i * ld + j] != 0.0;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) out[i] = sh[0] + sh[1] + sh[2] + sh[3];
}

// Row-block column ranges of the sorted K, the (row block, chunk) work list
// and, per column chunk, the row blocks whose range covers it.
tg_status build_band(tg_ctx* c) {
    int32_t *d_first = nullptr, *d_last = nullptr;
    TG_TRY(dev_alloc(c, &d_first, (size_t)c->n));
    TG_TRY(dev_alloc(c, &d_last, (size_t)c->n));
    hipLaunchKernelGGL(row_nz_extent_kernel, dim3(c->n), dim3(256), 0, c->stream, c->d_K, c->n,
                       c->ld, d_first, d_last);
    TG_CHECK_LAUNCH(c);
    std::vector<int32_t> first(c->n), last(c->n);
    TG_HIP(c, hipMemcpyAsync(first.data(), d_first, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
    TG_HIP(c, hipMemcpyAsync(last.data(), d_last, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
    TG_HIP(c, hipStreamSynchronize(c->stream));
    dev_free(&d_first);
    dev_free(&d_last);
    c->h_rb_col.assign(c->n_rowblk, int2{0, 0});
    std::vector<int2> rb_chunk(c->n_rowblk, int2{0, -1});
    std::vector<int2> work;
    std::vector<std::vector<int32_t>> cover(c->n_colchk);
    c->band_elems = 0;
    for (int32_t r = 0; r < c->n_rowblk; ++r) {
        int32_t lo = 0x7fffffff, hi = -1;
        const int64_t p1 = std::min<int64_t>(c->n, (int64_t)(r + 1) * kPlainRows);
        for (int64_t p = (int64_t)r * kPlainRows; p < p1; ++p)
            if (last[p] >= 0) {
                lo = std::min(lo, first[p]);
                hi = std::max(hi, last[p]);
            }
        if (hi < 0) continue;
        // whole sweeps: the kernel reads 128-column segments
        const int32_t c0 = lo / kPlainCols * kPlainCols;
        const int32_t c1 = std::min<int64_t>(c->ld, (hi / kPlainCols + 1) * kPlainCols);
        c->h_rb_col[r] = int2{c0, c1};
        rb_chunk[r] = int2{c0 / kPlainChunk, (c1 - 1) / kPlainChunk};
        for (int32_t ch = rb_chunk[r].x; ch <= rb_chunk[r].y; ++ch) {
            work.push_back(int2{r, ch});
            cover[ch].push_back(r);
        }
        c->band_elems += (p1 - (int64_t)r * kPlainRows) * (int64_t)(c1 - c0);
    }
    std::vector<int32_t> cc_off(c->n_colchk + 1, 0), cc_rows;
    for (int32_t ch = 0; ch < c->n_colchk; ++ch) {
        cc_off[ch + 1] = cc_off[ch] + (int32_t)cover[ch].size();
        cc_rows.insert(cc_rows.end(), cover[ch].begin(), cover[ch].end());
    }
    if (cc_rows.empty()) cc_rows.push_back(0);
    if (work.empty()) work.push_back(int2{0, 0});
    c->n_work = (int32_t)work.size();
    TG_TRY(dev_alloc(c, &c->d_rb_col, (size_t)c->n_rowblk));
    TG_TRY(dev_alloc(c, &c->d_rb_chunk, (size_t)c->n_rowblk));
    TG_TRY(dev_alloc(c, &c->d_work, work.size()));
    TG_TRY(dev_alloc(c, &c->d_cc_off, cc_off.size()));
    TG_TRY(dev_alloc(c, &c->d_cc_rows, cc_rows.size()));
    TG_HIP(c, hipMemcpyAsync(c->d_rb_col, c->h_rb_col.data(), sizeof(int2) * c->n_rowblk, hipMemcpyHostToDevice, c->stream));
    TG_HIP(c, hipMemcpyAsync(c->d_rb_chunk, rb_chunk.data(), sizeof(int2) * c->n_rowblk, hipMemcpyHostToDevice, c->stream));
    TG_HIP(c, hipMemcpyAsync(c->d_work, work.data(), sizeof(int2) * work.size(), hipMemcpyHostToDevice, c->stream));
    TG_HIP(c, hipMemcpyAsync(c->d_cc_off, cc_off.data(), sizeof(int32_t) * cc_off.size(), hipMemcpyHostToDevice, c->stream));
    TG_HIP(c, hipMemcpyAsync(c->d_cc_rows, cc_rows.data(), sizeof(int32_t) * cc_rows.size(), hipMemcpyHostToDevice, c->stream));
    TG_HIP(c, hipStreamSynchronize(c->stream));
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

        // 5. K (dense FP64, row stride ld, zero padding)
        TG_TRY(dev_alloc(c, &c->d_K, (size_t)c->n * c->ld));
        hipLaunchKernelGGL(kernel_entries, dim3(ceil_div(c->ld, kKThreads), c->n), dim3(kKThreads),
                           sizeof(double) * c->maxw, c->stream, c->d_R, c->d_Q, c->d_off, c->d_len,
                           c->maxw, c->qw, c->m_lo, c->nm, c->n, c->ld, c->d_perm, c->g * c->g,
                           c->d_K);
        TG_CHECK_LAUNCH(c);

        // 6. band of the sorted K and the work list of the one-pass kernel
        c->n_rowblk = (int32_t)ceil_div(c->n, kPlainRows);
        c->n_colchk = (int32_t)ceil_div(c->ld, kPlainChunk);
        TG_TRY(build_band(c));
        TG_TRY(dev_alloc(c, &c->d_upart, (size_t)c->n_colchk * c->n));
        TG_TRY(dev_alloc(c, &c->d_vpart, (size_t)c->n_rowblk * c->ld));
        // 7. K 1 and K^T 1 through the same one-pass kernel (f = 1)
        TG_TRY(dev_alloc(c, &c->d_rk, c->n));
        TG_TRY(dev_alloc(c, &c->d_ck, c->n));
        TG_TRY(dev_alloc(c, &c->d_rk_s, c->n));
        TG_TRY(dev_alloc(c, &c->d_ck_s, c->n));
        {
            std::vector<double> ones(c->n, 1.0);
            TG_HIP(c, hipMemcpyAsync(c->d_Q, ones.data(), sizeof(double) * c->n,
                                     hipMemcpyHostToDevice, c->stream));
            TG_TRY(plain_products_device(c, c->d_K, c->d_Q, c->d_rk, c->d_ck));
            hipLaunchKernelGGL(gather_kernel, dim3(ceil_div(c->n, 256)), dim3(256), 0, c->stream,
                               c->d_rk, c->d_perm, c->n, c->d_rk_s);
            hipLaunchKernelGGL(gather_kernel, dim3(ceil_div(c->n, 256)), dim3(256), 0, c->stream,
                               c->d_ck, c->d_perm, c->n, c->d_ck_s);
            TG_CHECK_LAUNCH(c);
            TG_HIP(c, hipStreamSynchronize(c->stream));
        }
        {
            unsigned long long* d_cnt = nullptr;
            TG_TRY(dev_alloc(c, &d_cnt, (size_t)c->n));
            hipLaunchKernelGGL(nnz_rows_kernel, dim3(c->n), dim3(256), 0, c->stream, c->d_K, c->n,
                               c->ld, d_cnt);
            TG_CHECK_LAUNCH(c);
            std::vector<unsigned long long> cnt(c->n);
            TG_HIP(c, hipMemcpyAsync(cnt.data(), d_cnt, sizeof(unsigned long long) * c->n,
                                     hipMemcpyDeviceToHost, c->stream));
            TG_HIP(c, hipStreamSynchronize(c->stream));
            dev_free(&d_cnt);
            c->nnz = 0;
            for (auto v : cnt) c->nnz += (int64_t)v;
        }
        TG_HIP(c, hipStreamSynchronize(c->stream));
        // the projection/convolution scratch is only needed for assembly
        dev_free(&c->d_Q);
        // 8. the FP64 band of K of the plain evaluation
        TG_TRY(stack_prepare(c));
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
    // K is stored in energy order: gather natural rows/columns through pinv
    std::vector<int32_t> pinv(c->n);
    for (int64_t p = 0; p < c->n; ++p) pinv[c->h_perm[p]] = (int32_t)p;
    int32_t* d_pinv = nullptr;
    double* d_out = nullptr;
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(nrows, (int64_t)(1ll << 27) / c->n));
    TG_TRY(dev_alloc(c, &d_pinv, (size_t)c->n));
    tg_status s = dev_alloc(c, &d_out, (size_t)chunk * c->n);
    if (s != TG_OK) {
        dev_free(&d_pinv);
        return s;
    }
    hipError_t e = hipMemcpyAsync(d_pinv, pinv.data(), sizeof(int32_t) * c->n, hipMemcpyHostToDevice,
                                  c->stream);
    for (int64_t r0 = 0; r0 < nrows && e == hipSuccess; r0 += chunk) {
        const int64_t nr = std::min(chunk, nrows - r0);
        hipLaunchKernelGGL(copy_rows_kernel, dim3(ceil_div(c->n, 256), nr), dim3(256), 0, c->stream,
                           c->d_K, c->n, c->ld, d_pinv, row0 + r0, d_out);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(out + r0 * c->n, d_out, sizeof(double) * nr * c->n,
                               hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    dev_free(&d_out);
    dev_free(&d_pinv);
    if (e != hipSuccess) return set_error(c, TG_ERR_HIP, "copy rows: %s", hipGetErrorString(e));
    return TG_OK;
}

namespace tg {

// dK/domega0: the same projection windows with the sampled derivative table
// G'(x) = -x/sigma^2 G(x) (x = m dE + omega0), then K' 1 and K'^T 1.
tg_status assemble_domega(tg_ctx* c) {
    if (c->d_Kd) return TG_OK;
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
    TG_TRY(dev_alloc(c, &d_gd, gd.size()));
    TG_TRY(dev_alloc(c, &d_Qd, (size_t)c->n * c->qw));
    TG_HIP(c, hipMemcpyAsync(d_gd, gd.data(), sizeof(double) * gd.size(), hipMemcpyHostToDevice,
                             c->stream));
    hipLaunchKernelGGL(conv_kernel, dim3(ceil_div(c->qw, 128), c->n), dim3(128), 0, c->stream,
                       c->d_R, c->d_len, c->maxw, d_gd, c->nm, c->qw, c->n, d_Qd);
    TG_CHECK_LAUNCH(c);
    TG_TRY(dev_alloc(c, &c->d_Kd, (size_t)c->n * c->ld));
    hipLaunchKernelGGL(kernel_entries, dim3(ceil_div(c->ld, kKThreads), c->n), dim3(kKThreads),
                       sizeof(double) * c->maxw, c->stream, c->d_R, d_Qd, c->d_off, c->d_len,
                       c->maxw, c->qw, c->m_lo, c->nm, c->n, c->ld, c->d_perm, c->g * c->g,
                       c->d_Kd);
    TG_CHECK_LAUNCH(c);
    TG_TRY(dev_alloc(c, &c->d_rkd, c->n));
    TG_TRY(dev_alloc(c, &c->d_ckd, c->n));
    std::vector<double> ones(c->n, 1.0);
    TG_HIP(c, hipMemcpyAsync(d_Qd, ones.data(), sizeof(double) * c->n, hipMemcpyHostToDevice,
                             c->stream));
    TG_TRY(plain_products_device(c, c->d_Kd, d_Qd, c->d_rkd, c->d_ckd));
    TG_HIP(c, hipStreamSynchronize(c->stream));
    dev_free(&d_Qd);
    dev_free(&d_gd);
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
